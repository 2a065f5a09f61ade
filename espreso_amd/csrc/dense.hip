// Batched dense assembly of the local dual operators on FP64 MFMA.
//
// For every subdomain (interior DOFs first, boundary DOFs = DOFs touched by B1
// after them) the host drives, tile step by tile step:
//   K_ii = L L^T                       (block-banded right-looking Cholesky)
//   [Y | y_f] = L^-1 [K_ib | f_i]      (blocked forward substitution, zero
//                                       leading tiles of each column skipped)
//   S = K_bb - Y^T Y                   (SYRK; this is the Dirichlet S_D)
//   S + rho Q Q^T = L_s L_s^T          (Q = orth(R_b), rho = max diag S)
//   W = L_s^-1,  F = W^T W             (F = (K^+)_bb, the local dual operator)
// which is the reference's S_i = B_i K_i^+ B_i^T (clustercpu.cpp:11-40 via
// SparseSolverMKL.cpp:1207-1428) and Dirichlet S_D (clustercpu.cpp:167-333)
// in interface-DOF space, with a boundary-kernel generalised inverse.
//
// Every GEMM-shaped step runs through one 64x64 tile routine on
// v_mfma_f64_16x16x4f64: 256 threads = 4 waves in a 2x2 arrangement, each
// wave owning a 32x32 quarter (2x2 MFMA blocks), K staged through LDS in
// chunks of 16.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace esp {

typedef double d4 __attribute__((ext_vector_type(4)));

// C(MxN) = alpha * op(A)(MxK) * op(B)(KxN) + beta * C, M,N <= 64.
// op(A)(i,k) = TA ? A[k + i*lda] : A[i + k*lda];  op(B)(k,j) = TB ? B[j + k*ldb] : B[k + j*ldb].
// C may alias A or B when this tile is the only reader of the aliased tile:
// all reads finish before the first store.  The next K chunk is fetched into
// registers while the MFMAs of the current chunk run from LDS.
template <bool TA, bool TB>
__device__ __forceinline__ void tile_gemm(double *C, int ldc, const double *A, int lda, const double *B, int ldb,
                                          int M, int N, int K, double alpha, double beta) {
    __shared__ double As[16][TILE + 1];
    __shared__ double Bs[16][TILE + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    d4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    // element (row/col index, k index) handled by this thread in load slot `it`
    auto aidx = [&](int it, int &i, int &kk) {
        const int q = tid + 256 * it;
        if (!TA) { i = q & 63; kk = q >> 6; } else { kk = q & 15; i = q >> 4; }
    };
    auto bidx = [&](int it, int &j, int &kk) {
        const int q = tid + 256 * it;
        if (!TB) { kk = q & 15; j = q >> 4; } else { j = q & 63; kk = q >> 6; }
    };
    double ra[4], rb[4];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            int i, kk;
            aidx(it, i, kk);
            ra[it] = (i < M && k0 + kk < K) ? (TA ? A[(k0 + kk) + (size_t)i * lda] : A[i + (size_t)(k0 + kk) * lda]) : 0.0;
            int j;
            bidx(it, j, kk);
            rb[it] = (j < N && k0 + kk < K) ? (TB ? B[j + (size_t)(k0 + kk) * ldb] : B[(k0 + kk) + (size_t)j * ldb]) : 0.0;
        }
    };
    if (K > 0) fetch(0);
    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            int i, kk;
            aidx(it, i, kk);
            As[kk][i] = ra[it];
            int j;
            bidx(it, j, kk);
            Bs[kk][j] = rb[it];
        }
        __syncthreads();
        if (k0 + 16 < K) fetch(k0 + 16);
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
            double a0 = As[ks + (lane >> 4)][wi + (lane & 15)];
            double a1 = As[ks + (lane >> 4)][wi + 16 + (lane & 15)];
            double b0 = Bs[ks + (lane >> 4)][wj + (lane & 15)];
            double b1 = Bs[ks + (lane >> 4)][wj + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout of the 16x16 f64 MFMA: value v of lane l is row (l/16)+4v, column l%16.
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int i = wi + 16 * a + (lane >> 4) + 4 * v;
                int j = wj + 16 * b + (lane & 15);
                if (i < M && j < N) {
                    double *c = C + i + (size_t)j * ldc;
                    double r = alpha * acc[a][b][v];
                    if (beta != 0.0) r += beta * *c;
                    *c = r;
                }
            }
}

__device__ __forceinline__ void pick(const DenseDesc &d, int which, double *&A, int &n, int &band, int &nt,
                                     double *&B, int &ncols, int &ntc) {
    if (which == 0) {
        A = d.Kii; n = d.ni; band = d.band_ii; nt = d.ntii; B = d.Kib; ncols = d.m + 1; ntc = d.ntc;
    } else {
        A = d.S; n = d.m; band = d.ntm; nt = d.ntm; B = d.W; ncols = d.m; ntc = d.ntm;
    }
}

// ---- left-looking block Cholesky, step k ----
// diagonal tile: A_kk -= L_kw L_kw^T over the band window w, then factor the
// tile and invert its factor (one workgroup per subdomain)
__global__ void __launch_bounds__(256) k_potrf_diag(const DenseDesc *D, int which, int step) {
    const DenseDesc &d = D[blockIdx.x];
    double *A, *B;
    int n, band, nt, ncols, ntc;
    pick(d, which, A, n, band, nt, B, ncols, ntc);
    const int k0 = step * TILE;
    if (k0 >= n) return;
    const int nb = min(TILE, n - k0);
    const int j0 = max(0, step - band) * TILE;
    double *Akk = A + k0 + (size_t)k0 * n;
    if (k0 > j0) {
        const double *Lk = A + k0 + (size_t)j0 * n;
        tile_gemm<false, true>(Akk, n, Lk, n, Lk, n, nb, nb, k0 - j0, -1.0, 1.0);
        __syncthreads();
    }
    __shared__ double T[TILE][TILE + 1];
    __shared__ double xd[TILE];
    __shared__ int bad;
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    for (int q = tid; q < TILE * TILE; q += 256) {
        int i = q & 63, j = q >> 6;
        double v;
        if (i < nb && j < nb) v = (i >= j) ? Akk[i + (size_t)j * n] : 0.0;
        else v = (i == j) ? 1.0 : 0.0;
        T[i][j] = v;
    }
    __syncthreads();
    for (int j = 0; j < TILE; j++) {
        if (tid == 0) {
            double dd = T[j][j];
            if (!(dd > 0.0) || !isfinite(dd)) {
                if (!bad) bad = j + 1;
                dd = 1.0;
            }
            T[j][j] = sqrt(dd);
        }
        __syncthreads();
        if (tid > j && tid < TILE) T[tid][j] /= T[j][j];
        __syncthreads();
        for (int q = tid; q < TILE * TILE; q += 256) {
            int i = q & 63, c = q >> 6;
            if (c > j && i >= c) T[i][c] -= T[i][j] * T[c][j];
        }
        __syncthreads();
    }
    // X = T^-1 (lower): thread c computes column c; X[i][c] for i > c is kept
    // in the unused strict upper triangle at T[c][i], X[c][c] in xd[c]
    if (tid < TILE) {
        const int c = tid;
        const double dc = 1.0 / T[c][c];
        xd[c] = dc;
        for (int i = c + 1; i < TILE; i++) {
            double s = T[i][c] * dc;
            for (int k = c + 1; k < i; k++) s += T[i][k] * T[c][k];
            T[c][i] = -s / T[i][i];
        }
    }
    __syncthreads();
    if (tid == 0 && bad && d.info[0] == 0) d.info[0] = k0 + bad;
    double *Dinv = d.Dinv + (size_t)step * TILE * TILE;
    for (int q = tid; q < TILE * TILE; q += 256) {
        int i = q & 63, j = q >> 6;
        Dinv[q] = i > j ? T[j][i] : (i == j ? xd[i] : 0.0);
        if (i < nb && j < nb && i >= j) Akk[i + (size_t)j * n] = T[i][j];
    }
}

// column below the diagonal: A_ik = (A_ik - L_iw L_kw^T) inv(L_kk)^T for
// step < i <= step + band, window w = [max(0, i - band), step)
__global__ void __launch_bounds__(256) k_potrf_panel(const DenseDesc *D, int which, int step) {
    const DenseDesc &d = D[blockIdx.y];
    double *A, *B;
    int n, band, nt, ncols, ntc;
    pick(d, which, A, n, band, nt, B, ncols, ntc);
    const int i = step + 1 + blockIdx.x;
    if (i >= nt || i > step + band) return;
    const int k0 = step * TILE, i0 = i * TILE;
    const int nb = min(TILE, n - k0), mi = min(TILE, n - i0);
    const int j0 = max(0, i - band) * TILE;
    double *C = A + i0 + (size_t)k0 * n;
    if (k0 > j0) {
        tile_gemm<false, true>(C, n, A + i0 + (size_t)j0 * n, n, A + k0 + (size_t)j0 * n, n, mi, nb, k0 - j0, -1.0,
                               1.0);
        __syncthreads();
    }
    tile_gemm<false, true>(C, n, C, n, d.Dinv + (size_t)step * TILE * TILE, TILE, mi, nb, nb, 1.0, 0.0);
}

// ---- left-looking forward substitution X = L^-1 B, block row i ----
// X_iJ = inv(L_ii) (B_iJ - L_iw X_wJ), window w = [max(first nonzero tile of
// column tile J, i - band), i)
__global__ void __launch_bounds__(256) k_trsm_row(const DenseDesc *D, int which, int step) {
    const DenseDesc &d = D[blockIdx.y];
    double *A, *B;
    int n, band, nt, ncols, ntc;
    pick(d, which, A, n, band, nt, B, ncols, ntc);
    const int J = blockIdx.x, i = step;
    if (J >= ntc || i >= nt) return;
    const int cs = which == 0 ? d.colstart[J] : J;
    if (cs > i) return;
    const int i0 = i * TILE, c0 = J * TILE;
    const int mi = min(TILE, n - i0), nj = min(TILE, ncols - c0);
    const int j0 = max(cs, i - band) * TILE;
    double *C = B + i0 + (size_t)c0 * n;
    if (i0 > j0) {
        tile_gemm<false, false>(C, n, A + i0 + (size_t)j0 * n, n, B + j0 + (size_t)c0 * n, n, mi, nj, i0 - j0, -1.0,
                                1.0);
        __syncthreads();
    }
    tile_gemm<false, false>(C, n, d.Dinv + (size_t)i * TILE * TILE, TILE, C, n, mi, nj, mi, 1.0, 0.0);
}

// lower tiles of  S = S - Y^T Y  (which 0)  or  S = W^T W  (which 1)
__global__ void __launch_bounds__(256) k_syrk(const DenseDesc *D, int which) {
    const DenseDesc &d = D[blockIdx.y];
    const int q = blockIdx.x;
    int I = (int)((sqrtf(8.0f * q + 1.0f) - 1.0f) * 0.5f);
    while ((I + 1) * (I + 2) / 2 <= q) I++;
    while (I * (I + 1) / 2 > q) I--;
    const int J = q - I * (I + 1) / 2;
    if (I >= d.ntm) return;
    const int m = d.m, i0 = I * TILE, j0 = J * TILE;
    const int mi = min(TILE, m - i0), mj = min(TILE, m - j0);
    double *C = d.S + i0 + (size_t)j0 * m;
    if (which == 0) {
        int ks = max(d.colstart[I], d.colstart[J]) * TILE;
        if (ks >= d.ni) return;
        const double *Y = d.Kib;
        tile_gemm<true, false>(C, m, Y + ks + (size_t)i0 * d.ni, d.ni, Y + ks + (size_t)j0 * d.ni, d.ni, mi, mj,
                               d.ni - ks, -1.0, 1.0);
    } else {
        int ks = i0;  // W is lower triangular: rows < max(I,J)*64 of these columns vanish
        tile_gemm<true, false>(C, m, d.W + ks + (size_t)i0 * m, m, d.W + ks + (size_t)j0 * m, m, mi, mj, m - ks,
                               1.0, 0.0);
    }
}

__global__ void k_fill(double *p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_identity(const DenseDesc *D) {
    const DenseDesc &d = D[blockIdx.y];
    const int64_t mm = (int64_t)d.m * d.m;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < mm; q += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = q % d.m, j = q / d.m;
        d.W[q] = (i == j) ? 1.0 : 0.0;
    }
}

__device__ __forceinline__ double block_max(double v) {
    __shared__ double red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

__global__ void k_rho(const DenseDesc *D, double *rho) {
    const DenseDesc &d = D[blockIdx.x];
    double v = 0.0;
    for (int i = threadIdx.x; i < d.m; i += 256) v = fmax(v, d.S[i + (size_t)i * d.m]);
    v = block_max(v);
    if (threadIdx.x == 0) rho[blockIdx.x] = v;
}

__global__ void k_kernel_reg(const DenseDesc *D, const double *Q, const int64_t *q_off, const int *kdim,
                             const double *rho) {
    const DenseDesc &d = D[blockIdx.y];
    const int m = d.m, k = kdim[blockIdx.y];
    const double r = rho[blockIdx.y];
    const double *q = Q + q_off[blockIdx.y];
    const int64_t mm = (int64_t)m * m;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mm; e += (int64_t)gridDim.x * blockDim.x) {
        int i = (int)(e % m), j = (int)(e / m);
        if (i < j) continue;
        double s = 0.0;
        for (int a = 0; a < k; a++) s += q[i + (size_t)a * m] * q[j + (size_t)a * m];
        d.S[e] += r * s;
    }
}

// hf[j] = fb[j] - sum_r Y[r][j] y_f[r]; one wave per boundary DOF
__global__ void k_hf(const DenseDesc *D, const double *fb, double *hf, const int64_t *off) {
    const DenseDesc &d = D[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= d.m) return;
    const double *y = d.Kib + (size_t)j * d.ni;
    const double *yf = d.Kib + (size_t)d.m * d.ni;
    double s = 0.0;
    for (int r = lane; r < d.ni; r += 64) s += y[r] * yf[r];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) hf[off[blockIdx.y] + j] = fb[off[blockIdx.y] + j] - s;
}

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__global__ void k_absmax(const DenseDesc *D, float *scale) {
    const DenseDesc &d = D[blockIdx.x];
    double v = 0.0;
    const int64_t mm = (int64_t)d.m * d.m;
    for (int64_t e = threadIdx.x; e < mm; e += 256) {
        int i = (int)(e % d.m), j = (int)(e / d.m);
        if (i >= j) v = fmax(v, fabs(d.S[e]));
    }
    v = block_max(v);
    if (threadIdx.x == 0) scale[blockIdx.x] = v > 0.0 ? (float)v : 1.0f;
}

// strip t, columns c = tR + blockIdx.y*256 + tid
template <int R, typename T>
__global__ void k_pack(const DenseDesc *D, T *pool, const int64_t *off, const float *scale) {
    const DenseDesc &d = D[blockIdx.z];
    const int m = d.m;
    const int t = blockIdx.x;
    const int r0 = t * R;
    if (r0 >= m) return;
    const int c = r0 + blockIdx.y * 256 + threadIdx.x;
    if (c >= m) return;
    const double inv = scale ? 1.0 / (double)scale[blockIdx.z] : 1.0;
    constexpr int VEC = 16 / (int)sizeof(T);
    T *base = pool + off[blockIdx.z];
#pragma unroll 4
    for (int r = 0; r < R; r++) {
        T *dst = base + strip_index(m, R, VEC, t, r, c);
        const int row = r0 + r;
        double v = 0.0;
        if (row < m && row <= c) {
            v = d.S[c + (size_t)row * m];
            if (row == c) v *= 0.5;
        }
        v *= inv;
        if constexpr (sizeof(T) == 8) *dst = (T)v;
        else if constexpr (sizeof(T) == 4) *dst = (T)(float)v;
        else *dst = to_bf16((float)v);
    }
}

__global__ void k_band_extract(const DenseDesc *D, double *band, const int64_t *band_off, const int *bw, double *yf,
                               const int64_t *yf_off) {
    const DenseDesc &d = D[blockIdx.y];
    const int w = bw[blockIdx.y] + 1;
    const int64_t total = (int64_t)d.ni * w;
    double *dst = band + band_off[blockIdx.y];
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int c = (int)(e / w), dd = (int)(e % w);
        dst[e] = (c + dd < d.ni) ? d.Kii[(c + dd) + (size_t)c * d.ni] : 0.0;
    }
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < d.ni; r += (int64_t)gridDim.x * blockDim.x)
        yf[yf_off[blockIdx.y] + r] = d.Kib[r + (size_t)d.m * d.ni];
}

template <int R, typename T>
__global__ void k_unpack(const T *pool, int64_t off, int m, float scale, double *out) {
    const int64_t mm = (int64_t)m * m;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mm; e += (int64_t)gridDim.x * blockDim.x) {
        int i = (int)(e % m), j = (int)(e / m);
        int r = min(i, j), c = max(i, j);
        int t = r / R;
        const T *src = pool + off + strip_index(m, R, 16 / (int)sizeof(T), t, r - t * R, c);
        double v;
        if constexpr (sizeof(T) == 8) v = (double)*src;
        else if constexpr (sizeof(T) == 4) v = (double)*src;
        else v = (double)__uint_as_float(((uint32_t)*src) << 16);
        v *= scale;
        if (i == j) v *= 2.0;
        out[e] = v;
    }
}

__global__ void k_scatter(const DenseDesc *D, const int *rowptr, const int64_t *rp_off, const int *colidx,
                          const double *vals, const int64_t *nz_off, const int *iperm, const int64_t *perm_off,
                          const double *f, const int *nloc) {
    const DenseDesc &d = D[blockIdx.y];
    const int b = blockIdx.y;
    const int n = nloc[b];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int *rp = rowptr + rp_off[b];
    const int *ci = colidx + nz_off[b];
    const double *kv = vals + nz_off[b];
    const int *ip = iperm + perm_off[b];
    const int ni = d.ni, m = d.m;
    const int pr = ip[r];
    for (int q = rp[r]; q < rp[r + 1]; q++) {
        const int pc = ip[ci[q]];
        const double v = kv[q];
        if (pr < ni) {
            if (pc < ni) d.Kii[pr + (size_t)pc * ni] = v;
            else d.Kib[pr + (size_t)(pc - ni) * ni] = v;
        } else if (pc >= ni) {
            d.S[(pr - ni) + (size_t)(pc - ni) * m] = v;
        }
    }
    if (pr < ni) d.Kib[pr + (size_t)m * ni] = f[perm_off[b] + r];
}

// ---- hybrid FETI: corner columns of F_i (lower triangle of S) ----
// FB0[j + q m] = sign_q F(j, j_q);  Fcc[q1 + q2 nq] = sign_q1 FB0[j_q1 + q2 m]
__global__ void k_extract_corners(const DenseDesc *D, const int *nq, const int64_t *qoff, const int *qj,
                                  const double *qsign, const int64_t *fb0off, const int64_t *fccoff, double *FB0,
                                  double *Fcc) {
    const DenseDesc &d = D[blockIdx.y];
    const int b = blockIdx.y, m = d.m, n = nq[b];
    const int *jq = qj + qoff[b];
    const double *sg = qsign + qoff[b];
    double *fb = FB0 + fb0off[b];
    const int64_t tot = (int64_t)m * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e % m), q = (int)(e / m);
        const int c = jq[q];
        const double v = j >= c ? d.S[j + (size_t)c * m] : d.S[c + (size_t)j * m];
        fb[e] = sg[q] * v;
    }
    if (blockIdx.x == 0) {
        double *fc = Fcc + fccoff[b];
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const int q1 = e % n, q2 = e / n;
            const int c1 = jq[q1], c2 = jq[q2];
            const double v = c1 >= c2 ? d.S[c1 + (size_t)c2 * m] : d.S[c2 + (size_t)c1 * m];
            fc[e] = sg[q1] * sg[q2] * v;
        }
    }
}

void launch_extract_corners(const DenseDesc *desc, int nb, const int *nq, const int64_t *qoff, const int *qj,
                            const double *qsign, const int64_t *fb0off, const int64_t *fccoff, double *FB0,
                            double *Fcc, int max_m, int max_nq, hipStream_t st) {
    unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(((int64_t)max_m * max_nq + 255) / 256, 1024));
    hipLaunchKernelGGL(k_extract_corners, dim3(g, nb), dim3(256), 0, st, desc, nq, qoff, qj, qsign, fb0off, fccoff,
                       FB0, Fcc);
}

// S(upper) = S(lower)^T for every descriptor
__global__ void k_symmetrize_desc(const DenseDesc *D) {
    const DenseDesc &d = D[blockIdx.y];
    const int m = d.m;
    const int64_t nn = (int64_t)m * m;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % m), j = (int)(e / m);
        if (i < j) d.S[e] = d.S[j + (int64_t)i * m];
    }
}

void launch_symmetrize_desc(const DenseDesc *desc, int nb, int max_m, hipStream_t st) {
    unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(((int64_t)max_m * max_m + 255) / 256, 1024));
    hipLaunchKernelGGL(k_symmetrize_desc, dim3(g, nb), dim3(256), 0, st, desc);
}

// ---- diagnostics: one MFMA with lane-indexed operands, one tile GEMM ----
__global__ void k_probe_mfma(const double *a, const double *b, double *out) {
    const int lane = threadIdx.x;
    d4 acc = d4{0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[lane], b[lane], acc, 0, 0, 0);
    for (int v = 0; v < 4; v++) out[lane * 4 + v] = acc[v];
}

__global__ void __launch_bounds__(256) k_probe_gemm(int ta, int tb, int M, int N, int K, const double *A, int lda,
                                                    const double *B, int ldb, double *C, int ldc, double alpha,
                                                    double beta) {
    if (!ta && !tb) tile_gemm<false, false>(C, ldc, A, lda, B, ldb, M, N, K, alpha, beta);
    else if (!ta && tb) tile_gemm<false, true>(C, ldc, A, lda, B, ldb, M, N, K, alpha, beta);
    else if (ta && !tb) tile_gemm<true, false>(C, ldc, A, lda, B, ldb, M, N, K, alpha, beta);
    else tile_gemm<true, true>(C, ldc, A, lda, B, ldb, M, N, K, alpha, beta);
}

void launch_probe_mfma(const double *a, const double *b, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_probe_mfma, dim3(1), dim3(64), 0, st, a, b, out);
}
void launch_probe_gemm(int ta, int tb, int M, int N, int K, const double *A, int lda, const double *B, int ldb,
                       double *C, int ldc, double alpha, double beta, hipStream_t st) {
    hipLaunchKernelGGL(k_probe_gemm, dim3(1), dim3(256), 0, st, ta, tb, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta);
}

// ------------------------------------------------------------ launchers ----
static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

void launch_fill(double *p, int64_t n, double v, hipStream_t st) {
    if (n <= 0) return;
    unsigned g = (unsigned)std::min<int64_t>(cdiv(n, 256), 65536);
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, st, p, n, v);
}

void launch_potrf_step(const DenseDesc *desc, int nb, int which, int step, int max_band, hipStream_t st) {
    hipLaunchKernelGGL(k_potrf_diag, dim3(nb), dim3(256), 0, st, desc, which, step);
    if (max_band <= 0) return;
    hipLaunchKernelGGL(k_potrf_panel, dim3(max_band, nb), dim3(256), 0, st, desc, which, step);
}

void launch_trsm_step(const DenseDesc *desc, int nb, int which, int step, int max_band, int max_ntc,
                      hipStream_t st) {
    (void)max_band;
    hipLaunchKernelGGL(k_trsm_row, dim3(max_ntc, nb), dim3(256), 0, st, desc, which, step);
}

void launch_syrk(const DenseDesc *desc, int nb, int which, int max_ntm, hipStream_t st) {
    hipLaunchKernelGGL(k_syrk, dim3(max_ntm * (max_ntm + 1) / 2, nb), dim3(256), 0, st, desc, which);
}

void launch_identity(const DenseDesc *desc, int nb, int max_m, hipStream_t st) {
    unsigned g = std::min<unsigned>(cdiv((int64_t)max_m * max_m, 256), 4096);
    hipLaunchKernelGGL(k_identity, dim3(g, nb), dim3(256), 0, st, desc);
}

void launch_kernel_reg(const DenseDesc *desc, int nb, const double *q, const int64_t *q_off, const int *kdim,
                       double *rho, int max_m, hipStream_t st) {
    hipLaunchKernelGGL(k_rho, dim3(nb), dim3(256), 0, st, desc, rho);
    unsigned g = std::min<unsigned>(cdiv((int64_t)max_m * max_m, 256), 4096);
    hipLaunchKernelGGL(k_kernel_reg, dim3(g, nb), dim3(256), 0, st, desc, q, q_off, kdim, rho);
}

void launch_hf(const DenseDesc *desc, int nb, const double *fb, double *hf, const int64_t *off, int max_m,
               hipStream_t st) {
    hipLaunchKernelGGL(k_hf, dim3(cdiv(max_m, 4), nb), dim3(256), 0, st, desc, fb, hf, off);
}

void launch_pack(const DenseDesc *desc, int nb, int prec, void *pool, const int64_t *off, float *scale,
                 int max_m, hipStream_t st) {
    const int R = strip_rows(prec);
    dim3 grid(cdiv(max_m, R), cdiv(max_m, 256), nb);
    if (prec == 0) {
        hipLaunchKernelGGL((k_pack<STRIP_F64, double>), grid, dim3(256), 0, st, desc, (double *)pool, off,
                           (const float *)nullptr);
    } else if (prec == 1) {
        hipLaunchKernelGGL((k_pack<STRIP_F32, float>), grid, dim3(256), 0, st, desc, (float *)pool, off,
                           (const float *)nullptr);
    } else {
        hipLaunchKernelGGL(k_absmax, dim3(nb), dim3(256), 0, st, desc, scale);
        hipLaunchKernelGGL((k_pack<STRIP_B16, uint16_t>), grid, dim3(256), 0, st, desc, (uint16_t *)pool, off,
                           (const float *)scale);
    }
}

void launch_band_extract(const DenseDesc *desc, int nb, double *band, const int64_t *band_off, const int *bw,
                         double *yf, const int64_t *yf_off, int max_ni, hipStream_t st) {
    unsigned g = std::min<unsigned>(cdiv((int64_t)max_ni * 64, 256), 4096);
    hipLaunchKernelGGL(k_band_extract, dim3(g, nb), dim3(256), 0, st, desc, band, band_off, bw, yf, yf_off);
}

void launch_unpack(const void *pool, int prec, int64_t off, int m, float scale, double *out, hipStream_t st) {
    unsigned g = std::min<unsigned>(cdiv((int64_t)m * m, 256), 8192);
    if (prec == 0)
        hipLaunchKernelGGL((k_unpack<STRIP_F64, double>), dim3(g), dim3(256), 0, st, (const double *)pool, off, m,
                           scale, out);
    else if (prec == 1)
        hipLaunchKernelGGL((k_unpack<STRIP_F32, float>), dim3(g), dim3(256), 0, st, (const float *)pool, off, m,
                           scale, out);
    else
        hipLaunchKernelGGL((k_unpack<STRIP_B16, uint16_t>), dim3(g), dim3(256), 0, st, (const uint16_t *)pool, off,
                           m, scale, out);
}

void launch_scatter(const DenseDesc *desc, int nb, const int *rowptr, const int64_t *rp_off, const int *colidx,
                    const double *vals, const int64_t *nz_off, const int *iperm, const int64_t *perm_off,
                    const double *f, const int *n, int max_n, hipStream_t st) {
    hipLaunchKernelGGL(k_scatter, dim3(cdiv(max_n, 256), nb), dim3(256), 0, st, desc, rowptr, rp_off, colidx, vals,
                       nz_off, iperm, perm_off, f, n);
}

}  // namespace esp
