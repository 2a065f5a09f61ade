// Dual-space kernels of the FETI iteration: the dual operator
// F x = sum_i B_i F_i B_i^T x (itersolvercpu.cpp:9-114), the preconditioners
// (itersolvercpu.cpp:520-602), the projector P = I - G^T (G G^T)^-1 G
// (itersolver.cpp:5541-5623), PCG vector updates and dot products
// (itersolver.cpp:1537-1684, 6427-6467), and the primal recovery
// (itersolver.cpp:372-431).
//
// F_i and S_D live in HBM in the packed-strip layout (see internal.h): one
// pass over the upper triangle applies both triangles.  A workgroup owns a
// contiguous range of strips of one subdomain; B_i^T x is gathered straight
// from lambda space into LDS, row and column contributions accumulate in LDS,
// and each workgroup writes one partial vector that a second pass sums and
// scatters back to lambda space in a fixed order (no global atomics).  The
// waves of a workgroup meet in 64-bit fixed-point accumulators (integer adds,
// so the order in which the waves arrive does not change a bit of the result:
// repeated applies are bitwise equal).
#include <hip/hip_runtime.h>

#include "dual.h"

namespace esp {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Acc { typedef float type; };
template <>
struct Acc<double> { typedef double type; };

template <typename T, typename A>
__device__ __forceinline__ A load_cvt(T v) { return (A)v; }
template <>
__device__ __forceinline__ float load_cvt<uint16_t, float>(uint16_t v) { return __uint_as_float(((uint32_t)v) << 16); }

// transpose-reduce R per-lane accumulators over the 64 lanes of a wave; on
// return lane l holds the total of row (l / (64/R)) in v[0]
template <int R, typename A>
__device__ __forceinline__ A wave_rows_reduce(A *v, int lane) {
    int n = R;
#pragma unroll
    for (int o = 32; o >= 64 / R; o >>= 1) {
        const bool up = lane & o;
        const int h = n >> 1;
#pragma unroll
        for (int i = 0; i < h; i++) {
            // select values, not addresses: a conditional over the two array
            // elements would index the array dynamically and push it to scratch
            const A lo = v[i], hi = v[i + h];
            const A send = up ? (A)lo : (A)hi;
            const A keep = up ? (A)hi : (A)lo;
            v[i] = keep + __shfl_xor(send, o, 64);
        }
        n = h;
    }
    A s = v[0];
#pragma unroll
    for (int o = 64 / R / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

// Fixed-point accumulation of the SYMV: value v enters the accumulators as
// rint(v 2^k).  With 2^e <= 2 ||F_s||_inf max|x| (the bound of every partial
// sum, doubled against the rounding of the stored values) and k = 60 - e,
// every accumulator stays below 2^60 in magnitude; one unit is 2^-59 of
// ||F_s||_inf max|x| or less, below the rounding of an FP64 sum of the same
// terms.
__device__ __forceinline__ unsigned long long to_fixed(double v, int k) {
    return (unsigned long long)(long long)__builtin_rint(ldexp(v, k));
}

template <typename T, int R>
__global__ void __launch_bounds__(256) k_symv(const SymvItem *__restrict__ items, const int *__restrict__ M,
                                              const int64_t *__restrict__ boff, const T *__restrict__ pool,
                                              const int64_t *__restrict__ poff, const float *__restrict__ fscale,
                                              const double *__restrict__ nrm, const int64_t *__restrict__ bl_ptr,
                                              const int *__restrict__ bl_lam, const double *__restrict__ coef,
                                              const double *__restrict__ x, int gather, double *__restrict__ ypart) {
    typedef typename Acc<T>::type A;
    extern __shared__ unsigned char smem[];
    __shared__ double wmax[4];
    const SymvItem it = items[blockIdx.x];
    const int s = it.sub, m = M[s];
    const int j0 = it.t0 * R, len = m - j0;
    unsigned long long *ys = (unsigned long long *)smem;  // len fixed-point sums
    A *xs = (A *)(ys + len);  // len + R entries: x of the columns, zero padding past m
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t b0 = boff[s] + j0;
    double xm = 0.0;
    for (int j = tid; j < len; j += 256) {
        double v;
        if (gather) {
            v = 0.0;
            for (int64_t q = bl_ptr[b0 + j]; q < bl_ptr[b0 + j + 1]; q++) v += coef[q] * x[bl_lam[q]];
        } else {
            v = x[b0 + j];
        }
        xs[j] = (A)v;
        ys[j] = 0ull;
        xm = fmax(xm, fabs(v));
    }
    if (tid < R) xs[len + tid] = (A)0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) xm = fmax(xm, __shfl_xor(xm, o, 64));
    if (lane == 0) wmax[wave] = xm;
    __syncthreads();
    xm = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    const double sc = fscale ? (double)fscale[s] : 1.0;
    const double bound = 2.0 * nrm[s] / sc * xm;
    // a non-finite x or operator gives NaN outputs (the fixed point would hide it)
    const bool finite = bound <= 1.7976931348623157e308;
    const int k = (finite && bound > 0.0) ? 60 - ilogb(bound) : 0;
    const T *base = pool + poff[s];
    for (int t = it.t0 + wave; t < it.t1; t += 4) {
        const int r0 = t * R;
        const int ncol = m - r0;
        const T *S = base + strip_offset(m, R, t);
        // the strip-row x: FP64 keeps it in registers (R = 16 values per lane);
        // BF16 (64 rows) keeps it in one register, row r in lane r, broadcast
        // with readlane (no LDS reads and waits in the chunk loop); FP32
        // (32 rows) reads it from LDS, which measured faster there
        constexpr bool XREG = sizeof(T) == 8, XLANE = sizeof(T) == 2;
        A racc[R], xr[XREG ? R : 1];
#pragma unroll
        for (int r = 0; r < R; r++) {
            racc[r] = (A)0;
            if constexpr (XREG) xr[r] = xs[r0 + r - j0];
        }
        const float xlane = XLANE ? (lane < R ? (float)xs[r0 + lane - j0] : 0.0f) : 0.0f;
        const A *xrow = xs + (r0 - j0);  // x of the strip rows (FP32: LDS broadcast reads)
        constexpr int VEC = 16 / (int)sizeof(T);
        constexpr int NV = R / VEC;
        for (int c0 = 0; c0 < ncol; c0 += 64) {
            const int w = min(64, ncol - c0);
            if (lane < w) {
                // chunk [R/VEC][w][VEC]: the q-th 16-byte load of lane l sits at (q*w + l)*VEC
                const u32x4 *src = (const u32x4 *)(S + (int64_t)c0 * R) + lane;
                u32x4 raw[NV];
#pragma unroll
                for (int q = 0; q < NV; q++) raw[q] = __builtin_nontemporal_load(src + (int64_t)q * w);
                const T *a = (const T *)raw;
                const int c = c0 + lane;
                const A xc = xs[r0 + c - j0];
                // four independent chains for the column sum (latency of the FMA chain)
                A col[4] = {(A)0, (A)0, (A)0, (A)0};
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const A av = load_cvt<T, A>(a[r]);
                    if constexpr (XREG) {
                        col[r & 3] = fma(av, xr[r], col[r & 3]);
                    } else if constexpr (XLANE) {
                        const A xrr = __builtin_bit_cast(
                            float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, xlane), r));
                        col[r & 3] = fma(av, xrr, col[r & 3]);
                    } else {
                        col[r & 3] = fma(av, xrow[r], col[r & 3]);
                    }
                    racc[r] = fma(av, xc, racc[r]);
                }
                atomicAdd(&ys[r0 + c - j0], to_fixed((double)((col[0] + col[1]) + (col[2] + col[3])), k));
            }
        }
        const A rs = wave_rows_reduce<R, A>(racc, lane);
        const int row = lane / (64 / R);
        if ((lane & (64 / R - 1)) == 0 && r0 + row < m) atomicAdd(&ys[r0 + row - j0], to_fixed((double)rs, k));
    }
    __syncthreads();
    for (int j = tid; j < len; j += 256)
        ypart[it.part_off + j] = finite ? ldexp((double)(long long)ys[j], -k) * sc : __builtin_nan("");
}

// yb[boff+j] = sum of the partials of the items of subdomain s covering j
__global__ void k_part_reduce(const SymvItem *items, const int *item0, const int *nitems, const int *M,
                              const int64_t *boff, int R, const double *ypart, double *yb) {
    const int s = blockIdx.x;
    const int m = M[s];
    const int i0 = item0[s], ni = nitems[s];
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        double v = 0.0;
        for (int q = 0; q < ni; q++) {
            const SymvItem it = items[i0 + q];
            const int j0 = it.t0 * R;
            if (j >= j0) v += ypart[it.part_off + j - j0];
        }
        yb[boff[s] + j] = v;
    }
}

// Block SYMV on MFMA (v_mfma_f64_16x16x4f64): up to 16 right-hand sides, the
// MFMA's N dimension (lanes of right-hand sides past nr carry zeros).  An
// item is at most BLOCK_STRIPS strips of one subdomain; wave w owns the
// 16-column blocks w, w+4, ... of the item's column range and visits every
// strip of the item that holds the block (the 16-column grid is common to all
// strips since strips start at multiples of 16):
//   column product  Dc[c][v] += sum_r A[r][c] x[r][v]   (A op: lane (c, row slot))
//   row product     Dr[r][v] += sum_c A[r][c] x[c][v]   (A op: lane (r, column slot))
// The 16-byte loads give the first operand layout; a per-wave LDS tile
// transposes the block for the second.  Column sums are final when the wave
// leaves the block (written straight to the item's partial, no atomics); row
// sums stay in registers per strip and meet in LDS at the end.  x and the
// partials are interleaved by right-hand side with stride nr.
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(256) k_symv_mfma(const SymvItem *__restrict__ items, const int *__restrict__ M,
                                                   const int64_t *__restrict__ boff, const double *__restrict__ pool,
                                                   const int64_t *__restrict__ poff, const double *__restrict__ xb,
                                                   int nr, double *__restrict__ ypart) {
    constexpr int R = STRIP_F64, S = BLOCK_STRIPS, TL = 18;  // TL: leading dimension of the transpose tile
    __shared__ double xr[S * R * 16], yr[S * R * 16];        // [item row][rhs], stride 16
    __shared__ double T[4][16 * TL];
    const SymvItem it = items[blockIdx.x];
    const int s = it.sub, m = M[s];
    const int ns = it.t1 - it.t0, r0i = it.t0 * R;
    const int nrow = min(ns * R, m - r0i);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const double *xs = xb + boff[s] * nr;
    for (int i = tid; i < S * R * 16; i += 256) {
        const int row = i >> 4, v = i & 15;
        xr[i] = (row < nrow && v < nr) ? xs[(int64_t)(r0i + row) * nr + v] : 0.0;
        yr[i] = 0.0;
    }
    __syncthreads();
    const double *bs = pool + poff[s];
    double *Tw = T[wave];
    v4d dr[S];
#pragma unroll
    for (int t = 0; t < S; t++) dr[t] = (v4d){0.0, 0.0, 0.0, 0.0};
    const int nblk = (m - r0i + 15) / 16;
    for (int cb = wave; cb < nblk; cb += 4) {
        const int c0 = r0i + 16 * cb;
        // row-product B operand: x[c0 + q + 4j][l16]
        double xc[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int c = c0 + q + 4 * j;
            xc[j] = (c < m && l16 < nr) ? xs[(int64_t)c * nr + l16] : 0.0;
        }
        v4d dc = {0.0, 0.0, 0.0, 0.0};
        const int tl = min(ns, cb + 1);  // strips holding the block (upper triangle)
#pragma unroll
        for (int tt = 0; tt < S; tt++) {
            if (tt < tl) {
                const int t = it.t0 + tt, r0 = t * R, ncol = m - r0;
                const int cl = c0 - r0;  // >= 0, multiple of 16
                const int k = cl >> 6, cc = (cl & 63) + l16;
                const int w = min(64, ncol - 64 * k);
                const u32x4 *chunk = (const u32x4 *)(bs + strip_offset(m, R, t) + (int64_t)64 * k * R);
                u32x4 p0 = {0u, 0u, 0u, 0u}, p1 = {0u, 0u, 0u, 0u};
                if (cc < w) {
                    p0 = __builtin_nontemporal_load(chunk + q * w + cc);        // rows 2q, 2q+1
                    p1 = __builtin_nontemporal_load(chunk + (q + 4) * w + cc);  // rows 2q+8, 2q+9
                }
                double a[4];
                __builtin_memcpy(&a[0], &p0, 16);
                __builtin_memcpy(&a[2], &p1, 16);
                // column product: row slot q of call (j, e) is row 2q + 8j + e
                const double *xrow = xr + (r0 - r0i) * 16;
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int e = 0; e < 2; e++)
                        dc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[2 * j + e], xrow[(2 * q + 8 * j + e) * 16 + l16],
                                                                  dc, 0, 0, 0);
                // transpose through the wave's tile: T[c][r]
                *(u32x4 *)(Tw + l16 * TL + 2 * q) = p0;
                *(u32x4 *)(Tw + l16 * TL + 2 * q + 8) = p1;
                wave_lds_sync();
                // row product: column slot q of call j is column q + 4j
#pragma unroll
                for (int j = 0; j < 4; j++)
                    dr[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(Tw[(q + 4 * j) * TL + l16], xc[j], dr[tt], 0, 0, 0);
                wave_lds_sync();
            }
        }
        // column sums of the block: lane holds D[c0 + q + 4i][l16]
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int c = c0 + q + 4 * i;
            if (c < m && l16 < nr) {
                const int j = c - r0i;
                if (j < nrow) yr[j * 16 + l16] = dc[i];  // the only column sum of j
                else ypart[(it.part_off + j) * nr + l16] = dc[i];
            }
        }
    }
    // row sums: lane holds D[strip row q + 4i][l16] per strip; the waves add
    // them in turn (a fixed order, so repeated applies are bitwise equal)
    __syncthreads();
    for (int w = 0; w < 4; w++) {
        if (wave == w) {
#pragma unroll
            for (int tt = 0; tt < S; tt++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int row = tt * R + q + 4 * i;
                    if (tt < ns && row < nrow && l16 < nr) yr[row * 16 + l16] += dr[tt][i];
                }
        }
        __syncthreads();
    }
    for (int i = tid; i < nrow * 16; i += 256) {
        const int j = i >> 4, v = i & 15;
        if (v < nr) ypart[(it.part_off + j) * nr + v] = yr[i];
    }
}

// yb[v * nbnd + boff + j] = sum of the block partials covering j (fixed order)
__global__ void k_part_reduce_mfma(const SymvItem *items, const int *item0, const int *nitems, const int *M,
                                   const int64_t *boff, int64_t nbnd, int nr, const double *ypart, double *yb) {
    const int s = blockIdx.x;
    const int m = M[s];
    const int i0 = item0[s], ni = nitems[s];
    for (int i = threadIdx.x; i < m * nr; i += blockDim.x) {
        const int j = i / nr, v = i % nr;
        double acc = 0.0;
        for (int q = 0; q < ni; q++) {
            const SymvItem it = items[i0 + q];
            const int j0 = it.t0 * STRIP_F64;
            if (j >= j0) acc += ypart[(it.part_off + j - j0) * nr + v];
        }
        yb[v * nbnd + boff[s] + j] = acc;
    }
}

// xb[j nr + v] = sum coef * x[v nlam + lambda], v < nr
__global__ void k_boundary_gather_multi(int64_t nb, int nr, const int64_t *bl_ptr, const int *bl_lam,
                                        const double *coef, const double *x, int64_t nlam, double *xb) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nb * nr;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i / nr;
        const int v = (int)(i % nr);
        double acc = 0.0;
        for (int64_t q = bl_ptr[j]; q < bl_ptr[j + 1]; q++) acc += coef[q] * x[v * nlam + bl_lam[q]];
        xb[i] = acc;
    }
}

// lambda-space y = sum_e coef_e * yb[b_e]
__global__ void k_lambda_gather(int nlam, const int *cnt, const int64_t *lb, const double *coef, const double *yb,
                                double *y) {
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < nlam; l += gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int e = 0; e < cnt[l]; e++) v += coef[2 * l + e] * yb[lb[2 * l + e]];
        y[l] = v;
    }
}

// boundary gather xb = B^T x with coefficients
__global__ void k_boundary_gather(int64_t nb, const int64_t *bl_ptr, const int *bl_lam, const double *coef,
                                  const double *x, double *xb) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nb; j += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int64_t q = bl_ptr[j]; q < bl_ptr[j + 1]; q++) v += coef[q] * x[bl_lam[q]];
        xb[j] = v;
    }
}

// boundary SpMV with K_bb (lumped preconditioner): yb = K_bb xb
__global__ void k_kbb(int64_t nb, const int64_t *ptr, const int64_t *col, const double *val, const double *xb,
                      double *yb) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nb; j += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int64_t q = ptr[j]; q < ptr[j + 1]; q++) v += val[q] * xb[col[q]];
        yb[j] = v;
    }
}

__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double r = 0.0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) r += red[w];
    return r;
}

// per subdomain s: out[off[s] + a] = -sum_j R_s[j, a] (B_s^T x)_j for a < k[s]
// (R_s: m_s x k_s column-major at R + roff[s]); one workgroup of 1024 threads
// per subdomain (enough loads in flight for a latency-bound gather).
// With the coarse basis this is G x = -(B R)^T x per subdomain.
constexpr int APPLY_G_THREADS = 1024;
__global__ void __launch_bounds__(APPLY_G_THREADS) k_apply_G(const int *M, const int64_t *boff, const int *kcols,
                                                             const int64_t *outoff, const int64_t *roff,
                                                             const double *R, const int64_t *bl_ptr,
                                                             const int *bl_lam, const double *sign, const double *x,
                                                             double *out) {
    constexpr int KB = 8;  // columns per pass
    constexpr int NW = APPLY_G_THREADS / 64;
    __shared__ double red[KB][NW];
    const int s = blockIdx.x, m = M[s], k = kcols[s];
    const int64_t b0 = boff[s];
    const double *Rs = R + roff[s];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int a0 = 0; a0 < k; a0 += KB) {
        const int kb = min(KB, k - a0);
        double acc[KB];
#pragma unroll
        for (int a = 0; a < KB; a++) acc[a] = 0.0;
        for (int j = threadIdx.x; j < m; j += APPLY_G_THREADS) {
            double xb = 0.0;
            for (int64_t q = bl_ptr[b0 + j]; q < bl_ptr[b0 + j + 1]; q++) xb += sign[q] * x[bl_lam[q]];
#pragma unroll
            for (int a = 0; a < KB; a++)
                if (a < kb) acc[a] += Rs[j + (int64_t)(a0 + a) * m] * xb;
        }
#pragma unroll
        for (int a = 0; a < KB; a++) {
            double v = acc[a];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) red[a][w] = v;
        }
        __syncthreads();
        if (threadIdx.x < kb) {
            double v = 0.0;
            for (int q = 0; q < NW; q++) v += red[threadIdx.x][q];
            out[outoff[s] + a0 + threadIdx.x] = -v;
        }
        __syncthreads();
    }
}

// d[coff[b] + a] = sum over the subdomains of coarse block b of their partials
__global__ void k_block_sum(int nblk, const int *blk_ptr, const int *blk_sub, const int64_t *subpart_off,
                            const int *ck, const int *coff, const double *part, double *d) {
    // one thread per (block, column); blocks hold at most 64 columns
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nblk * 64; e += gridDim.x * blockDim.x) {
        const int b = e >> 6, a = e & 63;
        const int q0 = blk_ptr[b], q1 = blk_ptr[b + 1];
        const int k = ck[blk_sub[q0]];
        if (a >= k) continue;
        double v = 0.0;
        for (int q = q0; q < q1; q++) v += part[subpart_off[blk_sub[q]] + a];
        d[coff[b] + a] = v;
    }
}

// e = GGinv d (symmetric, column access)
__global__ void __launch_bounds__(256) k_coarse_gemv(int nc, const double *Ginv, const double *d, double *e) {
    __shared__ double red[4];
    const int i = blockIdx.x;
    const double *col = Ginv + (int64_t)i * nc;
    double v = 0.0;
    for (int j = threadIdx.x; j < nc; j += 256) v += col[j] * d[j];
    v = block_sum(v, red);
    if (threadIdx.x == 0) e[i] = v;
}

// mode 0: y = x - G^T e ; mode 1: y = G^T e   (G^T e)_l = -sum_entries sign Rt_s[j,:] e_blk(s)
__global__ void k_project_finish(int nlam, const int *cnt, const int64_t *lb, const int *lsub, const double *sign,
                                 const int64_t *boff, const int *M, const int *ck, const int *cblk, const int *coff,
                                 const int64_t *rtoff, const double *Rt, const double *e, const double *x, double *y,
                                 int mode) {
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < nlam; l += gridDim.x * blockDim.x) {
        double g = 0.0;
        for (int q = 0; q < cnt[l]; q++) {
            const int s = lsub[2 * l + q];
            const int64_t j = lb[2 * l + q] - boff[s];
            const int m = M[s];
            const double *R = Rt + rtoff[s];
            const double *es = e + coff[cblk[s]];
            double acc = 0.0;
            for (int a = 0; a < ck[s]; a++) acc += R[j + (int64_t)a * m] * es[a];
            g -= sign[2 * l + q] * acc;
        }
        y[l] = mode == 0 ? x[l] - g : g;
    }
}

// ---- hybrid FETI: cluster correction of K^+ (cluster.cpp:796-916) ----
// g0 = B0 y_b, t = F0^-1 g0, alpha = Sa^-1 (G0 t - e0), lambda0 = F0^-1 (g0 - G0^T alpha) = t - H alpha
// with H = F0^-1 G0^T precomputed at setup.

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// t = F0^-1 g0: 32 rows per workgroup (8 per wave, accumulated together for
// memory-level parallelism), each row a dot over a column of the symmetric
// F0^-1; g0 is gathered from the SYMV output.
__global__ void __launch_bounds__(256) k_cluster_t(const ClusterDev C, const double *yb, double *t) {
    extern __shared__ double g0[];
    const int c = blockIdx.y;
    const int n0 = C.n0[c];
    const int rb = blockIdx.x * 32;
    if (rb >= n0) return;  // uniform over the workgroup
    const int64_t ro = C.rowoff[c];
    for (int r = threadIdx.x; r < n0; r += 256) {
        double v = 0.0;
        for (int64_t e = C.re_ptr[ro + r]; e < C.re_ptr[ro + r + 1]; e++) v += C.re_c[e] * yb[C.re_b[e]];
        g0[r] = v;
    }
    __syncthreads();
    const double *F = C.F0inv + C.f0off[c];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = rb + w * 8;
    double v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = 0.0;
    for (int q = lane; q < n0; q += 64) {
        const double gq = g0[q];
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (r0 + i < n0) v[i] += F[q + (int64_t)(r0 + i) * n0] * gq;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const double sv = wave_sum(v[i]);
        if (lane == 0 && r0 + i < n0) t[ro + r0 + i] = sv;
    }
}

// one workgroup per cluster: alpha = Sa^-1 (G0 t - e0), lambda0 = t - H alpha (in place over t)
__global__ void __launch_bounds__(1024) k_cluster_alpha(const ClusterDev C, const double *e0s, double *lam0,
                                                        double *alphaK) {
    extern __shared__ double sh[];
    const int c = blockIdx.x;
    const int n0 = C.n0[c], kt = C.kt[c];
    double *t2 = sh, *al = sh + C.max_kt;
    const int64_t ro = C.rowoff[c];
    const double *t = lam0 + ro;
    const double *G0T = C.G0T + C.g0off[c];  // n0 x kt
    const double *e0 = e0s + C.eoff[c];      // cluster-contiguous subdomain partials
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int a = w; a < kt; a += 16) {
        double v = 0.0;
        for (int r = lane; r < n0; r += 64) v += G0T[r + (int64_t)a * n0] * t[r];
        v = wave_sum(v);
        if (lane == 0) t2[a] = v - e0[a];
    }
    __syncthreads();
    const double *Sai = C.Sainv + C.saoff[c];
    for (int a = threadIdx.x; a < kt; a += 1024) {
        double v = 0.0;
        for (int b = 0; b < kt; b++) v += Sai[a + (int64_t)b * kt] * t2[b];
        al[a] = v;
        alphaK[C.eoff[c] + a] = v;
    }
    __syncthreads();
    const double *H = C.H + C.g0off[c];  // n0 x kt
    for (int r = threadIdx.x; r < n0; r += 1024) {
        double v = t[r];
        for (int a = 0; a < kt; a++) v -= H[r + (int64_t)a * n0] * al[a];
        lam0[ro + r] = v;
    }
}

// ---- hybrid FETI set-up on the device (cluster.cpp:1293-2016) ----
// F0 = sum over the cluster's subdomains of their corner blocks; one workgroup
// per cluster, subdomains in order (rows of one subdomain are distinct).
__global__ void __launch_bounds__(256) k_f0_assemble(const ClusterDev C, double *F0) {
    const int c = blockIdx.x;
    const int n0 = C.n0[c];
    double *F = F0 + C.f0off[c];
    const int64_t nn = (int64_t)n0 * n0;
    for (int64_t e = threadIdx.x; e < nn; e += 256) F[e] = 0.0;
    __syncthreads();
    const int64_t ro = C.rowoff[c];
    for (int s = C.sfirst[c]; s < C.sfirst[c] + C.scount[c]; s++) {
        const int nq = C.nq[s];
        const int *qr = C.qrow + C.qoff[s];
        const double *Fc = C.Fcc + C.fccoff[s];
        for (int e = threadIdx.x; e < nq * nq; e += 256) {
            const int q1 = e % nq, q2 = e / nq;
            F[(qr[q1] - ro) + (qr[q2] - ro) * (int64_t)n0] += Fc[e];
        }
        __syncthreads();
    }
}

// H(:, a) = F0^-1 G0^T(:, a): grid (kt, clusters), F0^-1 symmetric so rows are read along columns
__global__ void __launch_bounds__(256) k_cluster_h(const ClusterDev C) {
    extern __shared__ double g[];
    const int c = blockIdx.y, a = blockIdx.x;
    const int n0 = C.n0[c], kt = C.kt[c];
    if (a >= kt || n0 == 0) return;  // uniform
    const double *G0T = C.G0T + C.g0off[c] + (int64_t)a * n0;
    for (int r = threadIdx.x; r < n0; r += 256) g[r] = G0T[r];
    __syncthreads();
    const double *F = C.F0inv + C.f0off[c];
    double *H = C.H + C.g0off[c] + (int64_t)a * n0;
    for (int r = threadIdx.x; r < n0; r += 256) {
        double v = 0.0;
        for (int q = 0; q < n0; q++) v += F[r + (int64_t)q * n0] * g[q];
        H[r] = v;
    }
}

// Sa = G0 H + rho N N^T (rho = max diag of G0 H, 1 if that is not positive); one workgroup per cluster
__global__ void __launch_bounds__(256) k_cluster_sa(const ClusterDev C, double *Sa) {
    __shared__ double red[4];
    __shared__ double rho;
    const int c = blockIdx.x;
    const int n0 = C.n0[c], kt = C.kt[c], kn = C.kn[c];
    const double *G0T = C.G0T + C.g0off[c];
    const double *H = C.H + C.g0off[c];
    double *S = Sa + C.saoff[c];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = w; o < kt * kt; o += 4) {
        const int i = o % kt, j = o / kt;
        double v = 0.0;
        for (int r = lane; r < n0; r += 64) v += G0T[r + (int64_t)i * n0] * H[r + (int64_t)j * n0];
        v = wave_sum(v);
        if (lane == 0) S[o] = v;
    }
    __syncthreads();
    double m = 0.0;
    for (int i = threadIdx.x; i < kt; i += 256) m = fmax(m, S[i + (int64_t)i * kt]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (lane == 0) red[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        if (!(r > 0.0)) {
            // no corner rows (G0 empty): the scale of the cluster's own S_i, so
            // that the kernel term R Sa^-1 R^T stays within the size of K^+
            for (int q = 0; q < C.scount[c]; q++) r = fmax(r, C.rho_s[C.sfirst[c] + q]);
        }
        rho = r > 0.0 ? r : 1.0;
    }
    __syncthreads();
    const double *N = C.Nc + C.ncoff[c];  // kt x kn
    for (int o = threadIdx.x; o < kt * kt; o += 256) {
        const int i = o % kt, j = o / kt;
        double v = 0.0;
        for (int b = 0; b < kn; b++) v += N[i + (int64_t)b * kt] * N[j + (int64_t)b * kt];
        S[o] += rho * v;
    }
}

// y_b += -FB0_s lambda0 + addR * R_b alphaK; grid (row chunks of 256, subdomains);
// FB0 is C.FB0 (FP64) or C.FB0f (the low-precision F_i's columns)
template <typename T>
__global__ void __launch_bounds__(256) k_cluster_apply(const ClusterDev C, const T *FB0, const int *M,
                                                       const int64_t *boff, const int *kdim, const int64_t *rboff,
                                                       const double *Rb, const double *lambda0, const double *alphaK,
                                                       int addR, double *yb) {
    __shared__ double lam[256];
    const int s = blockIdx.y, m = M[s];
    if (blockIdx.x * 256 >= m) return;  // uniform
    const int nq = C.nq[s];
    const int64_t q0 = C.qoff[s];
    for (int q = threadIdx.x; q < nq; q += 256) lam[q] = lambda0[C.qrow[q0 + q]];
    __syncthreads();
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    const T *FB = FB0 + C.fb0off[s] + j;
    // eight independent chains: all loads of a group issue before their use
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int q = 0;
    for (; q + 8 <= nq; q += 8) {
        T f[8];
#pragma unroll
        for (int i = 0; i < 8; i++) f[i] = FB[(int64_t)(q + i) * m];
#pragma unroll
        for (int i = 0; i < 8; i++) acc[i] -= (double)f[i] * lam[q + i];
    }
    for (; q < nq; q++) acc[q & 7] -= (double)FB[(int64_t)q * m] * lam[q];
    double v = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    if (addR) {
        const double *R = Rb + rboff[s] + j;
        const double *al = alphaK + C.asub[s];
        for (int a = 0; a < kdim[s]; a++) v += R[(int64_t)a * m] * al[a];
    }
    yb[boff[s] + j] += v;
}

// FB0f = FB0 rounded to FP32.  With FP32 storage these are exactly the columns
// of the stored F_i, so the inner operator stays symmetric.  With BF16 storage
// they are not rounded to BF16: F0 = B0 F_i B0^T is built from the FP64 F_i,
// and the correction cancels the corner directions only when one of its two
// factors matches F0; a BF16-rounded pair leaves an indefinite error of the
// BF16 rounding on them (PCG then stalls).  grid (chunks, subdomains)
__global__ void __launch_bounds__(256) k_fb0_round(const ClusterDev C, const int *M) {
    const int s = blockIdx.y;
    const int64_t n = (int64_t)M[s] * C.nq[s];
    const double *src = C.FB0 + C.fb0off[s];
    float *dst = C.FB0f + C.fb0off[s];
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) dst[e] = (float)src[e];
}

void launch_fb0_round(const DualDev &D, const ClusterDev &C, hipStream_t st) {
    hipLaunchKernelGGL(k_fb0_round, dim3(64, D.nsub), dim3(256), 0, st, C, D.m);
}

// ---- the cluster correction for up to 8 cases per pass (block re-solve):
// F0^-1 and F_i B0^T are read once per pass instead of once per case.
// t_v = F0^-1 g0_v, g0_v gathered from the case's SYMV output; 4 rows per
// wave, kc <= 8 cases (g0 of the kc cases in LDS); grid (rows/16, clusters).
__global__ void __launch_bounds__(256) k_cluster_t_multi(const ClusterDev C, const double *yb, int64_t ystr,
                                                         int kc, double *t, int64_t tstr) {
    extern __shared__ double g0[];  // n0 x kc
    const int c = blockIdx.y;
    const int n0 = C.n0[c];
    const int rb = blockIdx.x * 16;
    if (rb >= n0) return;  // uniform over the workgroup
    const int64_t ro = C.rowoff[c];
    for (int e = threadIdx.x; e < n0 * kc; e += 256) {
        const int r = e / kc, v = e % kc;
        double g = 0.0;
        for (int64_t q = C.re_ptr[ro + r]; q < C.re_ptr[ro + r + 1]; q++) g += C.re_c[q] * yb[v * ystr + C.re_b[q]];
        g0[e] = g;
    }
    __syncthreads();
    const double *F = C.F0inv + C.f0off[c];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = rb + w * 4;
    double acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int v = 0; v < 8; v++) acc[i][v] = 0.0;
    for (int q = lane; q < n0; q += 64) {
        double f[4];
#pragma unroll
        for (int i = 0; i < 4; i++) f[i] = r0 + i < n0 ? F[q + (int64_t)(r0 + i) * n0] : 0.0;
#pragma unroll
        for (int v = 0; v < 8; v++) {
            if (v < kc) {
                const double gq = g0[q * kc + v];
#pragma unroll
                for (int i = 0; i < 4; i++) acc[i][v] += f[i] * gq;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int v = 0; v < 8; v++) {
            if (v < kc) {
                const double sv = wave_sum(acc[i][v]);
                if (lane == 0 && r0 + i < n0) t[v * tstr + ro + r0 + i] = sv;
            }
        }
}

// y_b,v += -FB0_s lambda0_v + addR * R_b alphaK_v for kc <= 8 cases;
// grid (row chunks of 256, subdomains)
__global__ void __launch_bounds__(256) k_cluster_apply_multi(const ClusterDev C, const int *M, const int64_t *boff,
                                                             const int *kdim, const int64_t *rboff, const double *Rb,
                                                             const double *lambda0, int64_t lstr,
                                                             const double *alphaK, int64_t astr, int addR, int kc,
                                                             double *yb, int64_t ystr) {
    __shared__ double lam[256 * 8];
    const int s = blockIdx.y, m = M[s];
    if (blockIdx.x * 256 >= m) return;  // uniform
    const int nq = C.nq[s];
    const int64_t q0 = C.qoff[s];
    for (int e = threadIdx.x; e < nq * kc; e += 256) {
        const int q = e / kc, v = e % kc;
        lam[e] = lambda0[v * lstr + C.qrow[q0 + q]];
    }
    __syncthreads();
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    const double *FB = C.FB0 + C.fb0off[s] + j;
    double acc[8];
#pragma unroll
    for (int v = 0; v < 8; v++) acc[v] = 0.0;
    for (int q = 0; q < nq; q++) {
        const double f = FB[(int64_t)q * m];
#pragma unroll
        for (int v = 0; v < 8; v++)
            if (v < kc) acc[v] -= f * lam[q * kc + v];
    }
    const double *R = Rb + rboff[s] + j;
    for (int v = 0; v < kc; v++) {
        double val = acc[v];
        if (addR) {
            const double *al = alphaK + v * astr + C.asub[s];
            for (int a = 0; a < kdim[s]; a++) val += R[(int64_t)a * m] * al[a];
        }
        yb[v * ystr + boff[s] + j] += val;
    }
}

// alpha_tot(s) = addK * alphaK(s) + N_s alpha_coarse(block(s))
__global__ void k_alpha_combine(int nsub, const int *kdim, const int *ck, const int *cblk, const int *coff,
                                const int64_t *asub, const int64_t *noff, const double *N, const double *alphaK,
                                int addK, const double *alpha_coarse, double *alpha_tot) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nsub; s += gridDim.x * blockDim.x) {
        const int k = kdim[s], kc = ck[s];
        const double *Ns = N + noff[s];  // k x kc column-major
        const double *ac = alpha_coarse + coff[cblk[s]];
        for (int a = 0; a < k; a++) {
            double v = addK ? alphaK[asub[s] + a] : 0.0;
            for (int b = 0; b < kc; b++) v += Ns[a + (int64_t)b * k] * ac[b];
            alpha_tot[asub[s] + a] = v;
        }
    }
}

// ---- vector kernels ----
__global__ void __launch_bounds__(256) k_dot_partial(const double *a, const double *b, int64_t n, double *partial) {
    __shared__ double red[4];
    double v = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v += a[i] * b[i];
    v = block_sum(v, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}
__global__ void __launch_bounds__(256) k_dot_final(const double *partial, int np, double *out) {
    __shared__ double red[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) v += partial[i];
    v = block_sum(v, red);
    if (threadIdx.x == 0) *out = v;
}

// p = y + beta p with beta = num/den (first: p = y)
__global__ void k_p_update(double *p, const double *y, int64_t n, const double *num, const double *den, int first) {
    const double beta = first ? 0.0 : (*num) / (*den);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = first ? y[i] : y[i] + beta * p[i];
}
// alpha = num/den; x += alpha p; r -= alpha Ap
__global__ void k_xr_update(double *x, double *r, const double *p, const double *Ap, int64_t n, const double *num,
                            const double *den) {
    const double alpha = (*num) / (*den);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        r[i] -= alpha * Ap[i];
    }
}
__global__ void k_axpby(double *y, const double *a, const double *b, int64_t n, double alpha, double beta) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = alpha * a[i] + beta * b[i];
}
__global__ void k_copy_scalar(const double *src, double *dst) { *dst = *src; }

__global__ void k_symmetrize(double *A, int n) {
    const int64_t nn = (int64_t)n * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn; e += (int64_t)gridDim.x * blockDim.x) {
        int i = (int)(e % n), j = (int)(e / n);
        if (i < j) A[e] = A[j + (int64_t)i * n];
    }
}


// ------------------------------------------------------------ launchers ----
static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }
static inline unsigned vgrid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, 256), 4096)); }

void launch_symv(const DualDev &D, bool sd, int prec, const double *x, int gather, bool scaled, hipStream_t st,
                 hipEvent_t ev0, hipEvent_t ev1) {
    if (sd) prec = D.sd_prec;
    const SymvItem *items = D.items[prec];
    const int n_items = D.n_items[prec];
    const double *coef = scaled ? D.bl_sscale : D.bl_sign;
    const double *nrm = sd ? D.sdnrm : D.fnrm;
    const size_t shm = symv_lds_bytes(D.max_m, prec);
    if (ev0) (void)hipEventRecord(ev0, st);
    if (prec == 0) {
        const double *pool = sd ? D.SD : D.Fd;
        const int64_t *off = sd ? D.sdoff : D.foff;
        hipLaunchKernelGGL((k_symv<double, STRIP_F64>), dim3(n_items), dim3(256), shm, st, items, D.m, D.boff, pool,
                           off, (const float *)nullptr, nrm, D.bl_ptr, D.bl_lam, coef, x, gather, D.ypart);
    } else if (prec == 1) {
        hipLaunchKernelGGL((k_symv<float, STRIP_F32>), dim3(n_items), dim3(256), shm, st, items, D.m, D.boff,
                           sd ? D.SDf : D.Ff, D.poff[1], (const float *)nullptr, nrm, D.bl_ptr, D.bl_lam, coef, x,
                           gather, D.ypart);
    } else {
        hipLaunchKernelGGL((k_symv<uint16_t, STRIP_B16>), dim3(n_items), dim3(256), shm, st, items, D.m, D.boff,
                           sd ? D.SDh : D.Fh, D.poff[2], (const float *)(sd ? D.sdscale : D.fscale), nrm, D.bl_ptr,
                           D.bl_lam, coef, x, gather, D.ypart);
    }
    if (ev1) (void)hipEventRecord(ev1, st);
    const int R = strip_rows(prec);
    hipLaunchKernelGGL(k_part_reduce, dim3(D.nsub), dim3(256), 0, st, items, D.item0[prec], D.nitems[prec], D.m,
                       D.boff, R, D.ypart, D.yb);
}

void launch_symv_mfma(const DualDev &D, const SymvBlock &B, bool sd, int nr, const double *x, bool scaled,
                      double *yb, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    if (nr < 1 || nr > 16) throw Error(ESPRESO_ERR_ARG, "block SYMV: 1 to 16 right-hand sides");
    const double *coef = scaled ? D.bl_sscale : D.bl_sign;
    const double *pool = sd ? D.SD : D.Fd;
    const int64_t *off = sd ? D.sdoff : D.foff;
    const unsigned gg = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(D.nbnd * nr, 256), 8192));
    hipLaunchKernelGGL(k_boundary_gather_multi, dim3(gg), dim3(256), 0, st, D.nbnd, nr, D.bl_ptr, D.bl_lam, coef, x,
                       (int64_t)D.nlam, B.xb);
    if (ev0) (void)hipEventRecord(ev0, st);
    hipLaunchKernelGGL(k_symv_mfma, dim3(B.n_items), dim3(256), 0, st, B.items, D.m, D.boff, pool, off, B.xb, nr,
                       B.ypart);
    if (ev1) (void)hipEventRecord(ev1, st);
    hipLaunchKernelGGL(k_part_reduce_mfma, dim3(D.nsub), dim3(256), 0, st, B.items, B.item0, B.nitems, D.m, D.boff,
                       D.nbnd, nr, B.ypart, yb);
}

void launch_lambda_gather(const DualDev &D, const double *yb, bool scaled, double *y, hipStream_t st) {
    hipLaunchKernelGGL(k_lambda_gather, dim3(vgrid(D.nlam)), dim3(256), 0, st, D.nlam, D.le_cnt, D.le_b,
                       scaled ? D.le_sscale : D.le_sign, yb, y);
}

void launch_boundary_gather(const DualDev &D, const double *x, bool scaled, double *xb, hipStream_t st) {
    hipLaunchKernelGGL(k_boundary_gather, dim3(vgrid(D.nbnd)), dim3(256), 0, st, D.nbnd, D.bl_ptr, D.bl_lam,
                       scaled ? D.bl_sscale : D.bl_sign, x, xb);
}

void launch_kbb(const DualDev &D, const double *xb, double *yb, hipStream_t st) {
    hipLaunchKernelGGL(k_kbb, dim3(vgrid(D.nbnd)), dim3(256), 0, st, D.nbnd, D.kbb_ptr, D.kbb_col, D.kbb_val, xb, yb);
}

void launch_project(const DualDev &D, const double *x, double *y, int mode, hipStream_t st) {
    // mode 0: y = P x;  mode 1: y = G^T GGinv x (x coarse);  mode 2: y = GGinv G x (coarse out)
    if (mode != 1) {
        hipLaunchKernelGGL(k_apply_G, dim3(D.nsub), dim3(APPLY_G_THREADS), 0, st, D.m, D.boff, D.ck, D.dsub_off, D.rtoff, D.Rt,
                           D.bl_ptr, D.bl_lam, D.bl_sign, x, D.dsub);
        hipLaunchKernelGGL(k_block_sum, dim3(vgrid((int64_t)D.nblk * 64)), dim3(256), 0, st, D.nblk, D.blk_ptr,
                           D.blk_sub, D.dsub_off, D.ck, D.coff, D.dsub, D.cd);
    }
    const double *din = mode == 1 ? x : D.cd;
    double *eout = mode == 2 ? y : D.ce;
    hipLaunchKernelGGL(k_coarse_gemv, dim3(D.ncoarse), dim3(256), 0, st, D.ncoarse, D.GGinv, din, eout);
    if (mode == 2) return;
    hipLaunchKernelGGL(k_project_finish, dim3(vgrid(D.nlam)), dim3(256), 0, st, D.nlam, D.le_cnt, D.le_b, D.le_sub,
                       D.le_sign, D.boff, D.m, D.ck, D.cblk, D.coff, D.rtoff, D.Rt, D.ce, x, y, mode);
}

void launch_subdomain_kernel_dots(const DualDev &D, const double *x, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_apply_G, dim3(D.nsub), dim3(APPLY_G_THREADS), 0, st, D.m, D.boff, D.kdim, D.asub, D.rboff, D.Rb,
                       D.bl_ptr, D.bl_lam, D.bl_sign, x, out);
}

void launch_cluster_correction(const DualDev &D, const ClusterDev &C, const double *e0s, double *yb, int addR,
                               hipStream_t st, bool lowp) {
    if (C.max_n0 > 0)
        hipLaunchKernelGGL(k_cluster_t, dim3((C.max_n0 + 31) / 32, C.ncl), dim3(256), (size_t)C.max_n0 * 8, st, C, yb,
                           C.lambda0);
    hipLaunchKernelGGL(k_cluster_alpha, dim3(C.ncl), dim3(1024), 2 * (size_t)std::max(C.max_kt, 1) * 8, st, C, e0s,
                       C.lambda0, C.alphaK);
    if (lowp && C.FB0f)
        hipLaunchKernelGGL(k_cluster_apply<float>, dim3((D.max_m + 255) / 256, D.nsub), dim3(256), 0, st, C,
                           (const float *)C.FB0f, D.m, D.boff, D.kdim, D.rboff, D.Rb, C.lambda0, C.alphaK, addR, yb);
    else
        hipLaunchKernelGGL(k_cluster_apply<double>, dim3((D.max_m + 255) / 256, D.nsub), dim3(256), 0, st, C,
                           (const double *)C.FB0, D.m, D.boff, D.kdim, D.rboff, D.Rb, C.lambda0, C.alphaK, addR, yb);
}

// t_v = F0^-1 B0 y_b,v for nr cases (case stride ystr in yb, tstr in t)
void launch_cluster_t_multi(const DualDev &D, const ClusterDev &C, int nr, const double *yb, int64_t ystr, double *t,
                            int64_t tstr, hipStream_t st) {
    if (C.max_n0 <= 0) return;
    const int kc_max = (int)std::max<int64_t>(1, std::min<int64_t>(8, 6144 / C.max_n0));
    for (int v0 = 0; v0 < nr; v0 += kc_max) {
        const int kc = std::min(kc_max, nr - v0);
        hipLaunchKernelGGL(k_cluster_t_multi, dim3((C.max_n0 + 15) / 16, C.ncl), dim3(256),
                           (size_t)C.max_n0 * kc * 8, st, C, yb + v0 * ystr, ystr, kc, t + v0 * tstr, tstr);
    }
    (void)D;
}

// alpha and lambda0 of one case from its t (in lam0) and e0s
void launch_cluster_alpha(const ClusterDev &C, const double *e0s, double *lam0, double *alphaK, hipStream_t st) {
    hipLaunchKernelGGL(k_cluster_alpha, dim3(C.ncl), dim3(1024), 2 * (size_t)std::max(C.max_kt, 1) * 8, st, C, e0s,
                       lam0, alphaK);
}

void launch_cluster_apply_multi(const DualDev &D, const ClusterDev &C, int nr, const double *lam0, int64_t lstr,
                                const double *alphaK, int64_t astr, int addR, double *yb, int64_t ystr,
                                hipStream_t st) {
    for (int v0 = 0; v0 < nr; v0 += 8) {
        const int kc = std::min(8, nr - v0);
        hipLaunchKernelGGL(k_cluster_apply_multi, dim3((D.max_m + 255) / 256, D.nsub), dim3(256), 0, st, C, D.m,
                           D.boff, D.kdim, D.rboff, D.Rb, lam0 + v0 * lstr, lstr, alphaK + v0 * astr, astr, addR, kc,
                           yb + v0 * ystr, ystr);
    }
}

void launch_cluster_setup_f0(const ClusterDev &C, hipStream_t st) {
    hipLaunchKernelGGL(k_f0_assemble, dim3(C.ncl), dim3(256), 0, st, C, C.F0inv);
}

void launch_cluster_setup_sa(const ClusterDev &C, double *Sa, hipStream_t st) {
    if (C.max_n0 > 0 && C.max_kt > 0)
        hipLaunchKernelGGL(k_cluster_h, dim3(C.max_kt, C.ncl), dim3(256), (size_t)C.max_n0 * 8, st, C);
    hipLaunchKernelGGL(k_cluster_sa, dim3(C.ncl), dim3(256), 0, st, C, Sa);
}

void launch_alpha_combine(const DualDev &D, const ClusterDev &C, bool hybrid, const double *alpha_coarse,
                          double *alpha_tot, hipStream_t st) {
    hipLaunchKernelGGL(k_alpha_combine, dim3(vgrid(D.nsub)), dim3(256), 0, st, D.nsub, D.kdim, D.ck, D.cblk, D.coff,
                       D.asub, D.noff, D.N, C.alphaK, hybrid ? 1 : 0, alpha_coarse, alpha_tot);
}

void launch_dot(const double *a, const double *b, int64_t n, double *partial, double *out, hipStream_t st) {
    const int np = 512;
    hipLaunchKernelGGL(k_dot_partial, dim3(np), dim3(256), 0, st, a, b, n, partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, st, partial, np, out);
}

void launch_p_update(double *p, const double *y, int64_t n, const double *num, const double *den, int first,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_p_update, dim3(vgrid(n)), dim3(256), 0, st, p, y, n, num, den, first);
}

void launch_xr_update(double *x, double *r, const double *p, const double *Ap, int64_t n, const double *num,
                      const double *den, hipStream_t st) {
    hipLaunchKernelGGL(k_xr_update, dim3(vgrid(n)), dim3(256), 0, st, x, r, p, Ap, n, num, den);
}

void launch_axpby(double *y, const double *a, const double *b, int64_t n, double alpha, double beta, hipStream_t st) {
    hipLaunchKernelGGL(k_axpby, dim3(vgrid(n)), dim3(256), 0, st, y, a, b, n, alpha, beta);
}

void launch_copy_scalar(const double *src, double *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(1), 0, st, src, dst);
}

void launch_symmetrize(double *A, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_symmetrize, dim3(vgrid((int64_t)n * n)), dim3(256), 0, st, A, n);
}

__global__ void k_copy(const double4 *__restrict__ src, double4 *__restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// read-only stream: each thread sums 4 independent 16-byte nontemporal loads
// per step; one double per workgroup keeps the loads alive
__global__ void __launch_bounds__(256) k_stream_read(const u32x4 *__restrict__ src, int64_t n, double *out) {
    __shared__ double red[4];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    unsigned acc = 0;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const u32x4 a = __builtin_nontemporal_load(src + i), b = __builtin_nontemporal_load(src + i + stride),
                    c = __builtin_nontemporal_load(src + i + 2 * stride),
                    d = __builtin_nontemporal_load(src + i + 3 * stride);
        acc ^= (a.x ^ b.y) ^ (c.z ^ d.w);
    }
    for (; i < n; i += stride) acc ^= __builtin_nontemporal_load(src + i).x;
    const double v = block_sum((double)acc, red);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

void launch_stream_read(const double *src, int64_t n, int wgs, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_stream_read, dim3(wgs), dim3(256), 0, st, (const u32x4 *)src, n / 2, out);
}

void launch_copy(const double *src, double *dst, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, st, (const double4 *)src, (double4 *)dst, n / 4);
}

// ---- banded interior solves (K_ii = L L^T, L a lower band of width bw stored
// by columns: L(c + d, c) at L[c w + d], w = bw + 1) for K vectors at once,
// interleaved z[r K + v].  Blocked by 64 columns: wave 0 solves the diagonal
// block with shuffles, then all threads apply the block to the rows it reaches.
constexpr int BAND_DL = 65;  // leading dimension of the staged diagonal block

// stage the lower band of columns [c0, c0 + nb) restricted to the diagonal
// block into Ds[c * BAND_DL + d] = L(c0 + c + d, c0 + c) (d <= bw, c + d < nb)
__device__ __forceinline__ void band_stage_diag(const double *__restrict__ L, int w, int bw, int c0, int nb,
                                                double *Ds) {
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        const int c = e >> 6, d = e & 63;
        Ds[c * BAND_DL + d] = (c < nb && d <= bw && c + d < nb) ? L[(int64_t)(c0 + c) * w + d] : 0.0;
    }
}

template <int K>
__device__ __forceinline__ void band_lower_solve(const double *__restrict__ L, int w, int bw, int ni, double *z,
                                                 double *Ds) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int c0 = 0; c0 < ni; c0 += 64) {
        const int nb = min(64, ni - c0);
        band_stage_diag(L, w, bw, c0, nb, Ds);
        __syncthreads();
        if (wave == 0) {
            double zi[K];
#pragma unroll
            for (int v = 0; v < K; v++) zi[v] = lane < nb ? z[(int64_t)(c0 + lane) * K + v] : 0.0;
            for (int c = 0; c < nb; c++) {
                const int d = lane - c;
                const double lv = d > 0 ? Ds[c * BAND_DL + d] : 0.0;
                const double dg = Ds[c * BAND_DL];
#pragma unroll
                for (int v = 0; v < K; v++) {
                    const double zc = __shfl(zi[v], c, 64) / dg;
                    zi[v] = lane == c ? zc : zi[v] - lv * zc;
                }
            }
            if (lane < nb)
#pragma unroll
                for (int v = 0; v < K; v++) z[(int64_t)(c0 + lane) * K + v] = zi[v];
        }
        __syncthreads();
        const int rend = min(ni, c0 + nb + bw);
        for (int r = c0 + nb + tid; r < rend; r += blockDim.x) {
            double acc[K];
#pragma unroll
            for (int v = 0; v < K; v++) acc[v] = 0.0;
#pragma unroll 8
            for (int c = max(c0, r - bw); c < c0 + nb; c++) {
                const double l = L[(int64_t)c * w + (r - c)];
#pragma unroll
                for (int v = 0; v < K; v++) acc[v] += l * z[(int64_t)c * K + v];
            }
#pragma unroll
            for (int v = 0; v < K; v++) z[(int64_t)r * K + v] -= acc[v];
        }
        __syncthreads();
    }
}

template <int K>
__device__ __forceinline__ void band_upper_solve(const double *__restrict__ L, int w, int bw, int ni, double *z,
                                                 double *Ds) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    for (int c0 = ((ni - 1) / 64) * 64; c0 >= 0; c0 -= 64) {
        const int nb = min(64, ni - c0);
        band_stage_diag(L, w, bw, c0, nb, Ds);
        // solved rows below the block: z[c] -= sum_{r >= c0 + nb, r - c <= bw} L(r, c) z[r]
        for (int cc = wave; cc < nb; cc += nw) {
            const int c = c0 + cc;
            const int rl = min(ni - 1, c + bw);
            double acc[K];
#pragma unroll
            for (int v = 0; v < K; v++) acc[v] = 0.0;
#pragma unroll 4
            for (int r = c0 + nb + lane; r <= rl; r += 64) {
                const double l = L[(int64_t)c * w + (r - c)];
#pragma unroll
                for (int v = 0; v < K; v++) acc[v] += l * z[(int64_t)r * K + v];
            }
#pragma unroll
            for (int v = 0; v < K; v++) acc[v] = wave_sum(acc[v]);
            if (lane == 0)
#pragma unroll
                for (int v = 0; v < K; v++) z[(int64_t)c * K + v] -= acc[v];
        }
        __syncthreads();
        if (wave == 0) {
            double zi[K];
#pragma unroll
            for (int v = 0; v < K; v++) zi[v] = lane < nb ? z[(int64_t)(c0 + lane) * K + v] : 0.0;
            for (int c = nb - 1; c >= 0; c--) {
                const int d = c - lane;  // L(c, i) of lanes i < c
                const double lv = (d > 0 && lane < nb) ? Ds[lane * BAND_DL + d] : 0.0;
                const double dg = Ds[c * BAND_DL];
#pragma unroll
                for (int v = 0; v < K; v++) {
                    const double zc = __shfl(zi[v], c, 64) / dg;
                    zi[v] = lane == c ? zc : zi[v] - lv * zc;
                }
            }
            if (lane < nb)
#pragma unroll
                for (int v = 0; v < K; v++) z[(int64_t)(c0 + lane) * K + v] = zi[v];
        }
        __syncthreads();
    }
}

// ---- new load vectors (K at once): y_f = L^-1 f_int (kept for the primal
// recovery) and hf = f_b - K_bi K_ii^-1 f_int, so that (K^+ f)_b = F_i hf.
// f, hf and yf hold the K vectors at strides fs, hs, ys; z is the interleaved
// work area (K values per interior row).  One workgroup per subdomain.
template <int K>
__global__ void __launch_bounds__(256) k_rhs_prepare(const PrimalDev P, const double *f, int64_t fstr, double *hf,
                                                     int64_t hstr, double *yf, int64_t ystr, double *work) {
    const int s = blockIdx.x;
    const int ni = P.ni[s], m = P.m[s], bw = P.bw[s];
    const int w = bw + 1;
    const double *L = P.band + P.band_off[s];
    const int *perm = P.perm + P.perm_off[s];
    const int *iperm = P.iperm + P.perm_off[s];
    const int *rp = P.rowptr + P.rp_off[s];
    const int *ci = P.colidx + P.nz_off[s];
    const double *kv = P.vals + P.nz_off[s];
    const double *fs = f + P.perm_off[s];
    double *z = work + P.work_off[s] * K;
    __shared__ double Ds[64 * BAND_DL];
    for (int i = threadIdx.x; i < ni * K; i += blockDim.x) z[i] = fs[(i % K) * fstr + perm[i / K]];
    __syncthreads();
    band_lower_solve<K>(L, w, bw, ni, z, Ds);
    for (int i = threadIdx.x; i < ni * K; i += blockDim.x) yf[(i % K) * ystr + P.yf_off[s] + i / K] = z[i];
    __syncthreads();
    band_upper_solve<K>(L, w, bw, ni, z, Ds);
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        const int lr = perm[ni + j];
        double v[K];
#pragma unroll
        for (int q = 0; q < K; q++) v[q] = fs[q * fstr + lr];
        for (int q = rp[lr]; q < rp[lr + 1]; q++) {
            const int pc = iperm[ci[q]];
            if (pc < ni)
#pragma unroll
                for (int r = 0; r < K; r++) v[r] -= kv[q] * z[(int64_t)pc * K + r];
        }
#pragma unroll
        for (int q = 0; q < K; q++) hf[q * hstr + P.boff[s] + j] = v[q];
    }
}

// ---- primal recovery (K at once): u_int = K_ii^-1 (f_int - K_ib u_b) via the
// banded factor, u = [u_int; u_b] + R alpha in local DOF order.
// First z = K_ib u_b on the interior rows (z[r K + t], r in factor order):
// workgroup (g, s) takes interior rows 64g..64g+63 of subdomain s, 16 lanes
// per CSR row reading consecutive entries (the same product in one thread
// per row would chain three dependent loads per entry).
constexpr int PRIMAL_ROWS = 64;
template <int K>
__global__ void __launch_bounds__(256) k_primal_kib(const PrimalDev P, const double *ub, int64_t bstr, double *work) {
    const int s = blockIdx.y, g = blockIdx.x;
    const int ni = P.ni[s];
    const int *perm = P.perm + P.perm_off[s];
    const int *iperm = P.iperm + P.perm_off[s];
    const int *rp = P.rowptr + P.rp_off[s];
    const int *ci = P.colidx + P.nz_off[s];
    const double *kv = P.vals + P.nz_off[s];
    const double *ubs = ub + P.boff[s];
    double *z = work + P.work_off[s] * K;
    const int grp = threadIdx.x >> 4, l = threadIdx.x & 15;
#pragma unroll
    for (int k = 0; k < PRIMAL_ROWS / 16; k++) {
        const int r = g * PRIMAL_ROWS + 16 * k + grp;
        if (r >= ni) break;  // uniform over the group
        const int lr = perm[r];
        double v[K];
#pragma unroll
        for (int t = 0; t < K; t++) v[t] = 0.0;
        for (int q = rp[lr] + l; q < rp[lr + 1]; q += 16) {
            const int pc = iperm[ci[q]];
            if (pc >= ni)
#pragma unroll
                for (int t = 0; t < K; t++) v[t] += kv[q] * ubs[t * bstr + pc - ni];
        }
#pragma unroll
        for (int t = 0; t < K; t++)
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) v[t] += __shfl_xor(v[t], o, 16);
        if (l < K) {
            double x = v[0];
#pragma unroll
            for (int t = 1; t < K; t++)
                if (l == t) x = v[t];
            z[(int64_t)r * K + l] = x;
        }
    }
}

// then, one workgroup per subdomain, the band solves and the assembly of u
template <int K>
__global__ void __launch_bounds__(256) k_primal_interior(const PrimalDev P, const double *ub, int64_t bstr,
                                                         const double *yf, int64_t ystr, const double *alpha,
                                                         int64_t astr, double *u, int64_t ustr, double *work) {
    const int s = blockIdx.x;
    const int ni = P.ni[s], n = P.n[s], bw = P.bw[s];
    const int w = bw + 1;
    const double *L = P.band + P.band_off[s];
    const int *perm = P.perm + P.perm_off[s];
    const double *ubs = ub + P.boff[s];
    double *z = work + P.work_off[s] * K;
    __shared__ double Ds[64 * BAND_DL];
    // z = K_ib u_b (k_primal_kib), then z <- y_f - L^-1 z = L^-1 (f_int - K_ib u_b)
    // and z <- L^-T z
    band_lower_solve<K>(L, w, bw, ni, z, Ds);
    for (int i = threadIdx.x; i < ni * K; i += blockDim.x) z[i] = yf[(i % K) * ystr + P.yf_off[s] + i / K] - z[i];
    __syncthreads();
    band_upper_solve<K>(L, w, bw, ni, z, Ds);
    const double *R = P.R + P.r_off[s];
    const int k = P.kdim[s];
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const int lr = perm[r];
#pragma unroll
        for (int q = 0; q < K; q++) {
            const double *al = alpha + q * astr + P.asub[s];
            double v = r < ni ? z[(int64_t)r * K + q] : ubs[q * bstr + r - ni];
            for (int a = 0; a < k; a++) v += R[lr + (int64_t)a * n] * al[a];
            u[q * ustr + P.perm_off[s] + lr] = v;
        }
    }
}

void launch_rhs_prepare(const PrimalDev &P, int nsub, const double *f, double *hf, hipStream_t st) {
    hipLaunchKernelGGL(k_rhs_prepare<1>, dim3(nsub), dim3(256), 0, st, P, f, (int64_t)0, hf, (int64_t)0, P.yf,
                       (int64_t)0, P.work);
}

// K vectors in chunks of 16, 8, 4, 2, 1; work holds 16 values per interior row
void launch_rhs_prepare_multi(const PrimalDev &P, int nsub, int nr, const double *f, int64_t fstr, double *hf,
                              int64_t hstr, double *yf, int64_t ystr, double *work, hipStream_t st) {
    for (int v0 = 0; v0 < nr;) {
        const int r = nr - v0;
        const int K = r >= 16 ? 16 : r >= 8 ? 8 : r >= 4 ? 4 : r >= 2 ? 2 : 1;
        const double *fv = f + v0 * fstr;
        double *hv = hf + v0 * hstr, *yv = yf + v0 * ystr;
#define ESP_RHS_CASE(KK)                                                                                          \
    case KK:                                                                                                      \
        hipLaunchKernelGGL(k_rhs_prepare<KK>, dim3(nsub), dim3(256), 0, st, P, fv, fstr, hv, hstr, yv, ystr, work); \
        break;
        switch (K) { ESP_RHS_CASE(16) ESP_RHS_CASE(8) ESP_RHS_CASE(4) ESP_RHS_CASE(2) ESP_RHS_CASE(1) }
#undef ESP_RHS_CASE
        v0 += K;
    }
}

// ---- structured grid on the device: K_i values (CSR of the grid builder)
// from the unit element matrix Ke and the element material scales.  One
// thread per CSR entry (row = local node x DOF, columns = the existing
// neighbours of the node in (z, y, x) offset order, DOFs inside); the entry
// sums scale_e * Ke over the elements holding both nodes in increasing (z, y,
// x) element order with separately rounded products and sums -- the order
// and rounding of the host builder's element loop, so the values agree bit
// for bit (structuralmechanics3d.cpp:353-509, physics.cpp:147-228).
__device__ __forceinline__ int hexa_corner(int cx, int cy, int cz) { return 4 * cz + (cy ? (cx ? 2 : 3) : (cx ? 1 : 0)); }

__global__ void __launch_bounds__(256) k_grid_stiffness(int nd, const int *box, int NE0, int NE1, const double *Ke,
                                                        const double *escale, const int *rowptr,
                                                        const int64_t *rp_off, const int64_t *nz_off, double *vals) {
    const int s = blockIdx.y;
    const int *bx = box + 6 * s;
    const int gx0 = bx[0], gy0 = bx[1], gz0 = bx[2], ex = bx[3], ey = bx[4], ez = bx[5];
    const int lx = ex + 1, ly = ey + 1, lz = ez + 1;
    const int n = lx * ly * lz * nd, ne = 8 * nd;
    const int *rp = rowptr + rp_off[s];
    const int nnz = rp[n];
    double *v = vals + nz_off[s];
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += gridDim.x * blockDim.x) {
#pragma clang fp contract(off)
        int lo = 0, hi = n;  // row r with rp[r] <= q < rp[r + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rp[mid] <= q) lo = mid;
            else hi = mid;
        }
        const int r = lo, ln = r / nd, i = r % nd;
        const int pos = q - rp[r], slot = pos / nd, j = pos % nd;
        const int x = ln % lx, y = (ln / lx) % ly, z = ln / (lx * ly);
        const int nxo = 1 + (x > 0) + (x < lx - 1), nyo = 1 + (y > 0) + (y < ly - 1);
        const int x2 = x + slot % nxo - (x > 0), y2 = y + (slot / nxo) % nyo - (y > 0),
                  z2 = z + slot / (nxo * nyo) - (z > 0);
        double acc = 0.0;
        for (int e2 = max(max(z, z2) - 1, 0); e2 <= min(min(z, z2), ez - 1); e2++)
            for (int e1 = max(max(y, y2) - 1, 0); e1 <= min(min(y, y2), ey - 1); e1++)
                for (int e0 = max(max(x, x2) - 1, 0); e0 <= min(min(x, x2), ex - 1); e0++) {
                    const int a = hexa_corner(x - e0, y - e1, z - e2), b = hexa_corner(x2 - e0, y2 - e1, z2 - e2);
                    const double sc = escale[(gx0 + e0) + (int64_t)NE0 * ((gy0 + e1) + (int64_t)NE1 * (gz0 + e2))];
                    const double t = sc * Ke[(nd * a + i) * ne + nd * b + j];  // rounded product (no FMA)
                    acc = acc + t;
                }
        v[q] = acc;
    }
}

__global__ void k_gather_values(int64_t n, const int64_t *pos, const double *src, double *dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[pos[i]];
}

// ---- structured grid: element sensitivities sens_e = -w_e^T K_e0 u_e ----
__global__ void __launch_bounds__(256) k_element_sensitivity(int NE0, int NE1, int NE2, int nd, const double *Ke0,
                                                             const double *u, const double *w, double *sens) {
    extern __shared__ double K[];
    const int ne = 8 * nd;
    for (int q = threadIdx.x; q < ne * ne; q += 256) K[q] = Ke0[q];
    __syncthreads();
    const int NN0 = NE0 + 1, NN1 = NE1 + 1;
    const int64_t nel = (int64_t)NE0 * NE1 * NE2;
    const int cx[8] = {0, 1, 1, 0, 0, 1, 1, 0}, cy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nel; e += (int64_t)gridDim.x * blockDim.x) {
        const int ex = (int)(e % NE0), ey = (int)((e / NE0) % NE1), ez = (int)(e / ((int64_t)NE0 * NE1));
        double ue[24], we[24];
        for (int a = 0; a < 8; a++) {
            const int64_t node = (ex + cx[a]) + (int64_t)NN0 * ((ey + cy[a]) + (int64_t)NN1 * (ez + cz[a]));
            for (int c = 0; c < nd; c++) {
                ue[a * nd + c] = u[node * nd + c];
                we[a * nd + c] = w[node * nd + c];
            }
        }
        double v = 0.0;
        for (int i = 0; i < ne; i++) {
            double t = 0.0;
            for (int j = 0; j < ne; j++) t += K[i * ne + j] * ue[j];
            v += we[i] * t;
        }
        sens[e] = -v;
    }
}

void launch_element_sensitivity(const int *NE, int nd, const double *Ke0, const double *u, const double *w,
                                double *sens, hipStream_t st) {
    const int64_t nel = (int64_t)NE[0] * NE[1] * NE[2];
    const int ne = 8 * nd;
    hipLaunchKernelGGL(k_element_sensitivity, dim3(vgrid(nel)), dim3(256), (size_t)ne * ne * 8, st, NE[0], NE[1],
                       NE[2], nd, Ke0, u, w, sens);
}

void launch_grid_stiffness(int nsub, int nd, const int *box, const int *NE, const double *Ke, const double *escale,
                           const int *rowptr, const int64_t *rp_off, const int64_t *nz_off, int max_nnz, double *vals,
                           hipStream_t st) {
    const unsigned gx = (unsigned)std::max(1, std::min((max_nnz + 255) / 256, 2048));
    hipLaunchKernelGGL(k_grid_stiffness, dim3(gx, nsub), dim3(256), 0, st, nd, box, NE[0], NE[1], Ke, escale, rowptr,
                       rp_off, nz_off, vals);
}

void launch_gather_values(int64_t n, const int64_t *pos, const double *src, double *dst, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_gather_values, dim3(vgrid(n)), dim3(256), 0, st, n, pos, src, dst);
}

void launch_primal(const PrimalDev &P, int nsub, const double *ub, double *u, hipStream_t st) {
    hipLaunchKernelGGL(k_primal_kib<1>, dim3((P.max_ni + PRIMAL_ROWS - 1) / PRIMAL_ROWS, nsub), dim3(256), 0, st, P,
                       ub, (int64_t)0, P.work);
    hipLaunchKernelGGL(k_primal_interior<1>, dim3(nsub), dim3(256), 0, st, P, ub, (int64_t)0, P.yf, (int64_t)0,
                       P.alpha, (int64_t)0, u, (int64_t)0, P.work);
}

void launch_primal_multi(const PrimalDev &P, int nsub, int nr, const double *ub, int64_t bstr, const double *yf,
                         int64_t ystr, const double *alpha, int64_t astr, double *u, int64_t ustr, double *work,
                         hipStream_t st) {
    for (int v0 = 0; v0 < nr;) {
        const int r = nr - v0;
        const int K = r >= 16 ? 16 : r >= 8 ? 8 : r >= 4 ? 4 : r >= 2 ? 2 : 1;
        const double *bv = ub + v0 * bstr, *yv = yf + v0 * ystr, *av = alpha + v0 * astr;
        double *uv = u + v0 * ustr;
#define ESP_PRIMAL_CASE(KK)                                                                                        \
    case KK:                                                                                                       \
        hipLaunchKernelGGL(k_primal_kib<KK>, dim3((P.max_ni + PRIMAL_ROWS - 1) / PRIMAL_ROWS, nsub), dim3(256), 0,  \
                           st, P, bv, bstr, work);                                                                 \
        hipLaunchKernelGGL(k_primal_interior<KK>, dim3(nsub), dim3(256), 0, st, P, bv, bstr, yv, ystr, av, astr, uv, \
                           ustr, work);                                                                            \
        break;
        switch (K) { ESP_PRIMAL_CASE(16) ESP_PRIMAL_CASE(8) ESP_PRIMAL_CASE(4) ESP_PRIMAL_CASE(2) ESP_PRIMAL_CASE(1) }
#undef ESP_PRIMAL_CASE
        v0 += K;
    }
}

// ---- SUPER_DIRICHLET preconditioner (clustercpu.cpp:243-314) ----
// S = K_bb - K_bi diag(K_ii)^-1 K_ib applied from the sparse blocks: first
// t = diag(K_nn)^-1 K_nb x_b for the interior DOFs next to the boundary,
// then y_b = K_bb x_b - K_bn t.
__global__ void k_superdir_near(int64_t n_near, const int64_t *ptr, const int64_t *col, const double *val,
                                const double *diag, const double *xb, double *t) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n_near; g += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int64_t q = ptr[g]; q < ptr[g + 1]; q++) v += val[q] * xb[col[q]];
        t[g] = v / diag[g];
    }
}

__global__ void k_superdir_bnd(int64_t nb, const int64_t *kptr, const int64_t *kcol, const double *kval,
                               const int64_t *ptr, const int64_t *col, const double *val, const double *xb,
                               const double *t, double *yb) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nb; j += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int64_t q = kptr[j]; q < kptr[j + 1]; q++) v += kval[q] * xb[kcol[q]];
        double w = 0.0;
        for (int64_t q = ptr[j]; q < ptr[j + 1]; q++) w += val[q] * t[col[q]];
        yb[j] = v - w;
    }
}

void launch_superdir(const DualDev &D, int64_t n_near, const int64_t *sib_ptr, const int64_t *sib_col,
                     const double *sib_val, const double *sdiag, const int64_t *sbi_ptr, const int64_t *sbi_col,
                     const double *sbi_val, double *t, const double *xb, double *yb, hipStream_t st) {
    if (n_near > 0)
        hipLaunchKernelGGL(k_superdir_near, dim3(vgrid(n_near)), dim3(256), 0, st, n_near, sib_ptr, sib_col, sib_val,
                           sdiag, xb, t);
    hipLaunchKernelGGL(k_superdir_bnd, dim3(vgrid(D.nbnd)), dim3(256), 0, st, D.nbnd, D.kbb_ptr, D.kbb_col, D.kbb_val,
                       sbi_ptr, sbi_col, sbi_val, xb, t, yb);
}

// ---- KKT report (itersolver.cpp:199-251) ----
// r = K_i u_i - (f_i - B1_i^T lambda) + B0_i^T lambda0 over the local DOFs of
// every subdomain; out[2s] = sum r^2, out[2s+1] = sum (f_i - B1_i^T lambda)^2.
// bt1 = B1^T lambda and bt0 = B0^T lambda0 (may be null) are boundary vectors;
// the boundary values of u are copied to ub (for B1 u).  Workgroup (g, s)
// takes rows 64g..64g+63 of subdomain s, 16 lanes per CSR row (the lanes of
// a group read consecutive entries of the row, 16 groups read 16 consecutive
// rows), and writes its two sums to part[2 (s G + g)]; k_kkt_sums adds the G
// partials of a subdomain in order (deterministic).
constexpr int KKT_ROWS = 64;
__global__ void __launch_bounds__(256) k_kkt_primal(const PrimalDev P, const double *u, const double *f,
                                                    const double *bt1, const double *bt0, double *ub, double *part) {
    __shared__ double red[4];
    const int s = blockIdx.y, g = blockIdx.x, G = gridDim.x;
    const int n = P.n[s], ni = P.ni[s];
    const int64_t uo = P.perm_off[s], ro = P.rp_off[s], zo = P.nz_off[s], bo = P.boff[s];
    const int grp = threadIdx.x >> 4, l = threadIdx.x & 15;
    double a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int k = 0; k < KKT_ROWS / 16; k++) {
        const int i = g * KKT_ROWS + 16 * k + grp;
        if (i >= n) break;  // uniform over the group
        double ku = 0.0;
        const int q1 = P.rowptr[ro + i + 1];
        for (int q = P.rowptr[ro + i] + l; q < q1; q += 16) ku += P.vals[zo + q] * u[uo + P.colidx[zo + q]];
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) ku += __shfl_xor(ku, o, 16);
        if (l == 0) {
            double t = f[uo + i], c0 = 0.0;
            const int pi = P.iperm[uo + i];
            if (pi >= ni) {
                const int64_t j = bo + pi - ni;
                t -= bt1[j];
                if (bt0) c0 = bt0[j];
                ub[j] = u[uo + i];
            }
            const double r = ku - t + c0;
            a1 += r * r;
            a2 += t * t;
        }
    }
    a1 = block_sum(a1, red);
    a2 = block_sum(a2, red);
    if (threadIdx.x == 0) {
        part[2 * ((int64_t)s * G + g)] = a1;
        part[2 * ((int64_t)s * G + g) + 1] = a2;
    }
}

__global__ void k_kkt_sums(int nsub, int G, const double *part, double *out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsub) return;
    double a1 = 0.0, a2 = 0.0;
    for (int g = 0; g < G; g++) {
        a1 += part[2 * ((int64_t)s * G + g)];
        a2 += part[2 * ((int64_t)s * G + g) + 1];
    }
    out[2 * s] = a1;
    out[2 * s + 1] = a2;
}

// bt0 = B0^T lambda0 on the boundary: one thread per subdomain walks its B0
// slots and their entries in order (a DOF may sit in several rows)
__global__ void k_b0t(int nsub, const int64_t *boff, const ClusterDev C, double *bt0) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsub) return;
    for (int64_t q = C.qoff[s]; q < C.qoff[s + 1]; q++) {
        const double l0 = C.lambda0[C.qrow[q]];
        for (int64_t e = C.qe_ptr[q]; e < C.qe_ptr[q + 1]; e++) bt0[boff[s] + C.qe_j[e]] += C.qe_c[e] * l0;
    }
}

int kkt_groups(int max_n) { return std::max(1, (max_n + KKT_ROWS - 1) / KKT_ROWS); }

void launch_kkt_primal(const PrimalDev &P, int nsub, int max_n, const double *u, const double *f, const double *bt1,
                       const double *bt0, double *ub, double *part, double *out, hipStream_t st) {
    const int G = kkt_groups(max_n);
    hipLaunchKernelGGL(k_kkt_primal, dim3(G, nsub), dim3(256), 0, st, P, u, f, bt1, bt0, ub, part);
    hipLaunchKernelGGL(k_kkt_sums, dim3((nsub + 255) / 256), dim3(256), 0, st, nsub, G, part, out);
}

void launch_b0t(const DualDev &D, const ClusterDev &C, double *bt0, hipStream_t st) {
    hipLaunchKernelGGL(k_b0t, dim3((D.nsub + 255) / 256), dim3(256), 0, st, D.nsub, D.boff, C, bt0);
}

// ---- output truncation (FETISolver.cpp:131-148) ----
__global__ void k_absmax_partial(const double *u, int64_t n, double *partial) {
    __shared__ double red[4];
    double m = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(u[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// D = 1 / precision / max(1, 10^ceil(log10(max|u|))); u <- trunc(D u) / D.
// The maximum starts at the smallest normal double, as the reference's does.
__global__ void k_truncate(double *u, int64_t n, const double *partial, int np, double precision) {
    __shared__ double dpl;
    if (threadIdx.x == 0) {
        double m = 2.2250738585072014e-308;
        for (int b = 0; b < np; b++) m = fmax(m, partial[b]);
        dpl = 1.0 / precision / fmax(1.0, pow(10.0, ceil(log10(m))));
    }
    __syncthreads();
    const double d = dpl;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        u[i] = trunc(d * u[i]) / d;
}

void launch_truncate(double *u, int64_t n, double precision, double *partial, hipStream_t st) {
    const int np = 256;
    hipLaunchKernelGGL(k_absmax_partial, dim3(np), dim3(256), 0, st, u, n, partial);
    hipLaunchKernelGGL(k_truncate, dim3(vgrid(n)), dim3(256), 0, st, u, n, (const double *)partial, np, precision);
}

}  // namespace esp
