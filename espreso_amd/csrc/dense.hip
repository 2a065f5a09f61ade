// Batched dense assembly of the local dual operators on FP64 MFMA.
//
// For every subdomain (interior DOFs first, boundary DOFs = DOFs touched by B1
// after them) the host drives, tile step by tile step:
//   K_ii = L L^T                       (block-banded right-looking Cholesky)
//   [Y | y_f] = L^-1 [K_ib | f_i]      (blocked forward substitution, zero
//                                       leading tiles of each column skipped)
//   S = K_bb - Y^T Y                   (SYRK writing -Y^T Y, then the CSR
//                                       entries of K_bb added; the Dirichlet S_D)
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

// Loads through the global address space: the operand pointers of the tile
// routine live in descriptors, so the compiler cannot prove them global and
// would emit flat loads, which also count in lgkmcnt; the wait for the LDS
// operand reads before the MFMAs would then wait for the prefetch of the next
// chunk as well.
__device__ __forceinline__ double gload(const double *p) {
    return *(const __attribute__((address_space(1))) double *)p;
}

// C(MxN) = alpha * op(A)(MxK) * op(B)(KxN) + beta * C, M,N <= 64.
// op(A)(i,k) = TA ? A[k + i*lda] : A[i + k*lda];  op(B)(k,j) = TB ? B[j + k*ldb] : B[k + j*ldb].
// C may alias A or B when this tile is the only reader of the aliased tile:
// all reads finish before the first store.  The next K chunk is fetched into
// registers while the MFMAs of the current chunk run from LDS, and stored into
// the second LDS buffer once they are issued (one barrier per chunk).
template <bool TA, bool TB>
__device__ __forceinline__ void tile_gemm(double *C, int ldc, const double *A, int lda, const double *B, int ldb,
                                          int M, int N, int K, double alpha, double beta) {
    // two LDS buffers: chunk c+1 is stored into one while the MFMAs read chunk c
    // from the other, so each 16-wide K chunk needs one workgroup barrier
    __shared__ double As[2][16][TILE + 1];
    __shared__ double Bs[2][16][TILE + 1];
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
    int fk = 0;  // K offset of the chunk held in ra/rb
    // out-of-range elements load element 0 of the operand (in bounds), with no
    // branch around the loads; store() replaces them by zero, so the wait for
    // the loads falls after the MFMAs of the chunk in between
    auto fetch = [&](int k0) {
        fk = k0;
#pragma unroll
        for (int it = 0; it < 4; it++) {
            int i, kk;
            aidx(it, i, kk);
            const bool va = i < M && k0 + kk < K;
            ra[it] = gload(A + (va ? (TA ? (k0 + kk) + (size_t)i * lda : i + (size_t)(k0 + kk) * lda) : 0));
            int j;
            bidx(it, j, kk);
            const bool vb = j < N && k0 + kk < K;
            rb[it] = gload(B + (vb ? (TB ? j + (size_t)(k0 + kk) * ldb : (k0 + kk) + (size_t)j * ldb) : 0));
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            int i, kk;
            aidx(it, i, kk);
            As[buf][kk][i] = (i < M && fk + kk < K) ? ra[it] : 0.0;
            int j;
            bidx(it, j, kk);
            Bs[buf][kk][j] = (j < N && fk + kk < K) ? rb[it] : 0.0;
        }
    };
    if (K > 0) {
        fetch(0);
        store(0);
        __syncthreads();
        if (16 < K) fetch(16);
    }
    for (int k0 = 0, cur = 0; k0 < K; k0 += 16, cur ^= 1) {
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
            double a0 = As[cur][ks + (lane >> 4)][wi + (lane & 15)];
            double a1 = As[cur][ks + (lane >> 4)][wi + 16 + (lane & 15)];
            double b0 = Bs[cur][ks + (lane >> 4)][wj + (lane & 15)];
            double b1 = Bs[cur][ks + (lane >> 4)][wj + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        // buffer cur^1 was last read in the previous chunk, before the barrier that ended it
        if (k0 + 16 < K) store(cur ^ 1);
        __syncthreads();
        if (k0 + 32 < K) fetch(k0 + 32);
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

// ---- fused tile steps: C = Dinv * (beta C - op(A) op(B)), the update kept in registers ----
// The 64 x 64 tile is split into column strips: wave w owns columns 16w..16w+15
// and all 64 rows (row blocks rb = 0..3), so the update R of its strip stays in
// the MFMA accumulators.  Value v of acc[rb] in lane l is R(16 rb + l/16 + 4 v,
// 16 w + l%16), which is the B operand of a 16x16x4 MFMA whose k-step covers
// rows 16 rb + 4 v + (0..3): the multiply by the staged diagonal inverse takes
// R straight from the accumulators, without a round trip of C through memory.
constexpr int FUSED_LDS = 2 * 2 * 16 * (TILE + 1);  // doubles: two operand buffers, or one padded 64 x 64 tile

// acc[rb] += op(A)(rows of block rb, :K) op(B)(:K, the wave's columns), M x N result, K > 0
template <bool TA, bool TB>
__device__ __forceinline__ void strip_gemm(d4 (&acc)[4], double *lds, const double *A, int lda, const double *B,
                                           int ldb, int M, int N, int K) {
    double(*As)[16][TILE + 1] = reinterpret_cast<double(*)[16][TILE + 1]>(lds);
    double(*Bs)[16][TILE + 1] = reinterpret_cast<double(*)[16][TILE + 1]>(lds + 2 * 16 * (TILE + 1));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    auto aidx = [&](int it, int &i, int &kk) {
        const int q = tid + 256 * it;
        if (!TA) { i = q & 63; kk = q >> 6; } else { kk = q & 15; i = q >> 4; }
    };
    auto bidx = [&](int it, int &j, int &kk) {
        const int q = tid + 256 * it;
        if (!TB) { kk = q & 15; j = q >> 4; } else { j = q & 63; kk = q >> 6; }
    };
    double ra[4], rb[4];
    int fk = 0;
    auto fetch = [&](int k0) {
        fk = k0;
#pragma unroll
        for (int it = 0; it < 4; it++) {
            int i, kk;
            aidx(it, i, kk);
            const bool va = i < M && k0 + kk < K;
            ra[it] = gload(A + (va ? (TA ? (k0 + kk) + (size_t)i * lda : i + (size_t)(k0 + kk) * lda) : 0));
            int j;
            bidx(it, j, kk);
            const bool vb = j < N && k0 + kk < K;
            rb[it] = gload(B + (vb ? (TB ? j + (size_t)(k0 + kk) * ldb : (k0 + kk) + (size_t)j * ldb) : 0));
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            int i, kk;
            aidx(it, i, kk);
            As[buf][kk][i] = (i < M && fk + kk < K) ? ra[it] : 0.0;
            int j;
            bidx(it, j, kk);
            Bs[buf][kk][j] = (j < N && fk + kk < K) ? rb[it] : 0.0;
        }
    };
    fetch(0);
    store(0);
    __syncthreads();
    if (16 < K) fetch(16);
    for (int k0 = 0, cur = 0; k0 < K; k0 += 16, cur ^= 1) {
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
            const double b = Bs[cur][ks + (lane >> 4)][16 * wave + (lane & 15)];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const double a = As[cur][ks + (lane >> 4)][16 * r + (lane & 15)];
                acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[r], 0, 0, 0);
            }
        }
        if (k0 + 16 < K) store(cur ^ 1);
        __syncthreads();  // after the last chunk: the operand buffers are free
        if (k0 + 32 < K) fetch(k0 + 32);
    }
}

// R = beta * C0 - acc on the rows < M, columns < N of the tile (zero elsewhere),
// then C = Dinv(:, :M) R.  TR: the tile's element (r, c) is C[c + r ldc] (the
// transposed step), else C[r + c ldc].  Dinv is 64 x 64, column-major.  The
// operand buffers must be free (no wave still reads them).
template <bool TR, bool KEEP = false>
__device__ __forceinline__ void strip_finish(d4 (&acc)[4], double *lds, double *C, int ldc, int M, int N,
                                             const double *Dinv, double beta, d4 *keep = nullptr) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = 16 * wave + (lane & 15);
    auto at = [&](int r) { return TR ? (size_t)c + (size_t)r * ldc : (size_t)r + (size_t)c * ldc; };
    // stage Dinv(i, k) at lds[k (TILE + 1) + i], columns k >= M zero; staged
    // before C is read, so that its registers are free again (occupancy)
#pragma unroll
    for (int h = 0; h < 2; h++) {
        double dv[8];
#pragma unroll
        for (int it = 0; it < 8; it++) dv[it] = gload(Dinv + tid + 256 * (8 * h + it));
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int e = tid + 256 * (8 * h + it), i = e & 63, k = e >> 6;
            lds[k * (TILE + 1) + i] = k < M ? dv[it] : 0.0;
        }
    }
    if (beta != 0.0) {
#pragma unroll
        for (int rbk = 0; rbk < 4; rbk++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = 16 * rbk + (lane >> 4) + 4 * v;
                const bool ok = r < M && c < N;
                const double c0 = gload(C + (ok ? at(r) : 0));
                acc[rbk][v] = ok ? beta * c0 - acc[rbk][v] : 0.0;
            }
    } else {
#pragma unroll
        for (int rbk = 0; rbk < 4; rbk++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = 16 * rbk + (lane >> 4) + 4 * v;
                acc[rbk][v] = (r < M && c < N) ? -acc[rbk][v] : 0.0;
            }
    }
    __syncthreads();
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
        d4 out = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int rbk = 0; rbk < 4; rbk++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double a = lds[(16 * rbk + 4 * v + (lane >> 4)) * (TILE + 1) + 16 * ob + (lane & 15)];
                out = __builtin_amdgcn_mfma_f64_16x16x4f64(a, acc[rbk][v], out, 0, 0, 0);
            }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = 16 * ob + (lane >> 4) + 4 * v;
            if (r < M && c < N) C[at(r)] = out[v];
        }
        if constexpr (KEEP) keep[ob] = out;
    }
}

// A_ii -= L L^T for the panel tile L (mi x 64, zero past its valid rows and
// columns) that strip_finish<true, true> just produced: lane value v of o[ob]
// is L(16 w + l%16, 16 ob + l/16 + 4v).  L goes through LDS (no re-read of
// the stored tile) and the 64 x 64 x 64 product runs on MFMA, wave w owning
// columns 16w..16w+15 of A_ii.  A_ii is written by this workgroup only.
__device__ __forceinline__ void diag_update(const d4 (&o)[4], double *lds, double *Aii, int lda, int mi) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __syncthreads();  // every wave has read the staged diagonal inverse
#pragma unroll
    for (int ob = 0; ob < 4; ob++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            lds[(16 * wave + (lane & 15)) * (TILE + 1) + 16 * ob + (lane >> 4) + 4 * v] = o[ob][v];
    __syncthreads();
    d4 acc[4] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int ks = 0; ks < TILE; ks += 4) {
        const double b = lds[(16 * wave + (lane & 15)) * (TILE + 1) + ks + (lane >> 4)];
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const double a = lds[(16 * rb + (lane & 15)) * (TILE + 1) + ks + (lane >> 4)];
            acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[rb], 0, 0, 0);
        }
    }
    const int q = 16 * wave + (lane & 15);
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int p = 16 * rb + (lane >> 4) + 4 * v;
            if (p < mi && q < mi) Aii[p + (size_t)q * lda] -= acc[rb][v];
        }
}

// ---- two column tiles per workgroup (forward substitution) ----
// The tiles (i, J) and (i, J+1) of one block row share the A operand (rows i of
// the factor) and the diagonal inverse: each wave owns 16 columns of both
// tiles.  A tile's B rows before its own window start (kfirst[s]) are read as
// zero: the other tile may start its window earlier.
constexpr int FUSED2_LDS = 2 * 16 * (TILE + 1) + 2 * 16 * (2 * TILE + 1);  // doubles

__device__ __forceinline__ void strip2_gemm(d4 (&acc)[2][4], double *lds, const double *A, int lda, const double *B,
                                            int ldb, int M, int N, int K, const int (&kfirst)[2]) {
    double(*As)[16][TILE + 1] = reinterpret_cast<double(*)[16][TILE + 1]>(lds);
    double(*Bs)[16][2 * TILE + 1] = reinterpret_cast<double(*)[16][2 * TILE + 1]>(lds + 2 * 16 * (TILE + 1));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double ra[4], rb[8];
    int fk = 0;
    auto bok = [&](int j, int k) { return j < N && k < K && k >= kfirst[j >> 6]; };
    auto fetch = [&](int k0) {
        fk = k0;
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const int q = tid + 256 * it, i = q & 63, kk = q >> 6;
            const bool va = i < M && k0 + kk < K;
            ra[it] = gload(A + (va ? i + (size_t)(k0 + kk) * lda : 0));
        }
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int q = tid + 256 * it, kk = q & 15, j = q >> 4;
            rb[it] = gload(B + (bok(j, k0 + kk) ? (k0 + kk) + (size_t)j * ldb : 0));
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 4; it++) {
            const int q = tid + 256 * it, i = q & 63, kk = q >> 6;
            As[buf][kk][i] = (i < M && fk + kk < K) ? ra[it] : 0.0;
        }
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int q = tid + 256 * it, kk = q & 15, j = q >> 4;
            Bs[buf][kk][j] = bok(j, fk + kk) ? rb[it] : 0.0;
        }
    };
    fetch(0);
    store(0);
    __syncthreads();
    if (16 < K) fetch(16);
    for (int k0 = 0, cur = 0; k0 < K; k0 += 16, cur ^= 1) {
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
            const int kr = ks + (lane >> 4);
            double a[4];
#pragma unroll
            for (int r = 0; r < 4; r++) a[r] = As[cur][kr][16 * r + (lane & 15)];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const double b = Bs[cur][kr][TILE * t + 16 * wave + (lane & 15)];
#pragma unroll
                for (int r = 0; r < 4; r++) acc[t][r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b, acc[t][r], 0, 0, 0);
            }
        }
        if (k0 + 16 < K) store(cur ^ 1);
        __syncthreads();  // after the last chunk: the operand buffers are free
        if (k0 + 32 < K) fetch(k0 + 32);
    }
}

// R_t = beta C0 - acc_t on rows < M, columns < N of tile t (columns 64 t..64 t+63
// of C, element (r, c) at C[r + c ldc]), then C = Dinv(:, :M) R_t for the valid tiles
__device__ __forceinline__ void strip2_finish(d4 (&acc)[2][4], double *lds, double *C, int ldc, int M, int N,
                                              const double *Dinv, double beta, const bool (&valid)[2]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        double dv[8];
#pragma unroll
        for (int it = 0; it < 8; it++) dv[it] = gload(Dinv + tid + 256 * (8 * h + it));
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int e = tid + 256 * (8 * h + it), i = e & 63, k = e >> 6;
            lds[k * (TILE + 1) + i] = k < M ? dv[it] : 0.0;
        }
    }
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int c = TILE * t + 16 * wave + (lane & 15);
#pragma unroll
        for (int rbk = 0; rbk < 4; rbk++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = 16 * rbk + (lane >> 4) + 4 * v;
                const bool ok = valid[t] && r < M && c < N;
                const double c0 = (beta != 0.0) ? gload(C + (ok ? (size_t)r + (size_t)c * ldc : 0)) : 0.0;
                acc[t][rbk][v] = ok ? beta * c0 - acc[t][rbk][v] : 0.0;
            }
    }
    __syncthreads();
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
        d4 out[2] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
        for (int rbk = 0; rbk < 4; rbk++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double a = lds[(16 * rbk + 4 * v + (lane >> 4)) * (TILE + 1) + 16 * ob + (lane & 15)];
#pragma unroll
                for (int t = 0; t < 2; t++) out[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, acc[t][rbk][v], out[t], 0, 0, 0);
            }
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int c = TILE * t + 16 * wave + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = 16 * ob + (lane >> 4) + 4 * v;
                if (valid[t] && r < M && c < N) C[(size_t)r + (size_t)c * ldc] = out[t][v];
            }
        }
    }
}

// Workgroups are dealt to the 8 XCDs (each with its own L2) round-robin by
// linear id, x fastest.  The tiles of one subdomain share their operand panels,
// so the linear id is remapped to keep runs of consecutive logical tiles (x
// fastest, then subdomain) on one XCD: physical p of a full round of 8 goes to
// logical (p % 8) * (G / 8) + p / 8; the tail past the last full round stays.
constexpr int NXCD = 8;
__device__ __forceinline__ void xcd_tile(int &bx, int &by) {
    const int gx = gridDim.x, G = gx * gridDim.y;
    const int p = blockIdx.x + gx * blockIdx.y;
    const int full = G - G % NXCD;
    const int L = p < full ? (p % NXCD) * (full / NXCD) + p / NXCD : p;
    bx = L % gx;
    by = L / gx;
}

// the factored matrix A (n x n, leading dimension lda) and the right-hand
// sides B (n x ncols, leading dimension n) of a stage
__device__ __forceinline__ void pick(const DenseDesc &d, int which, double *&A, int &lda, int &n, int &band, int &nt,
                                     double *&B, int &ncols, int &ntc) {
    if (which == 0) {
        A = d.Kii; lda = d.ldk; n = d.ni; band = d.band_ii; nt = d.ntii; B = d.Kib; ncols = d.m + 1; ntc = d.ntc;
    } else {
        A = d.S; lda = d.m; n = d.m; band = d.ntm; nt = d.ntm; B = d.W; ncols = d.m; ntc = d.ntm;
    }
}

// ---- block Cholesky, step k ----
// diagonal tile: factor the tile and invert its factor (one workgroup per
// subdomain).  Its updates by the block columns to its left, A_kk -= L_kj
// L_kj^T, were applied by the panel steps j that produced L_kj, so this
// latency-bound chain of launches carries no GEMM of its own.
__global__ void __launch_bounds__(256) k_potrf_diag(const DenseDesc *D, int which, int step) {
    const DenseDesc &d = D[blockIdx.x];
    double *A, *B;
    int lda, n, band, nt, ncols, ntc;
    pick(d, which, A, lda, n, band, nt, B, ncols, ntc);
    const int k0 = step * TILE;
    if (k0 >= n) return;
    const int nb = min(TILE, n - k0);
    double *Akk = A + k0 + (size_t)k0 * lda;
    __shared__ double T[TILE][TILE + 1];
    __shared__ double xd[TILE];
    __shared__ double tmp[3][16][17];
    __shared__ int bad;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) bad = 0;
    for (int q = tid; q < TILE * TILE; q += 256) {
        int i = q & 63, j = q >> 6;
        double v;
        if (i < nb && j < nb) v = (i >= j) ? Akk[i + (size_t)j * lda] : 0.0;
        else v = (i == j) ? 1.0 : 0.0;
        T[i][j] = v;
    }
    __syncthreads();
    // ---- T = L L^T, right-looking over four 16-column blocks ----
    for (int kb = 0; kb < 4; kb++) {
        const int c0 = kb * 16;
        if (wave == 0) {
            // diagonal block in registers: lane r < 16 holds row r
            const int r = lane;
            double v[16];
#pragma unroll
            for (int c = 0; c < 16; c++) v[c] = (r < 16 && c <= r) ? T[c0 + r][c0 + c] : 0.0;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                double piv = __shfl(v[j], j, 64);
                if (!(piv > 0.0) || !isfinite(piv)) {
                    if (lane == 0 && bad == 0) bad = c0 + j + 1;
                    piv = 1.0;
                }
                const double dd = sqrt(piv);
                if (r == j) v[j] = dd;
                else if (r > j) v[j] = v[j] / dd;
#pragma unroll
                for (int c = j + 1; c < 16; c++) {
                    const double lc = __shfl(v[j], c, 64);
                    if (r >= c) v[c] -= v[j] * lc;
                }
            }
            if (r < 16) {
#pragma unroll
                for (int c = 0; c < 16; c++)
                    if (c <= r) T[c0 + r][c0 + c] = v[c];
            }
        }
        __syncthreads();
        const int rest = TILE - c0 - 16;
        if (rest == 0) break;
        // panel rows below: L_rb = A_rb L_bb^-T, one thread per row
        if (tid < rest) {
            const int r = c0 + 16 + tid;
            double x[16];
#pragma unroll
            for (int c = 0; c < 16; c++) {
                double sv = T[r][c0 + c];
#pragma unroll
                for (int i = 0; i < c; i++) sv -= x[i] * T[c0 + c][c0 + i];
                x[c] = sv / T[c0 + c][c0 + c];
            }
#pragma unroll
            for (int c = 0; c < 16; c++) T[r][c0 + c] = x[c];
        }
        __syncthreads();
        // trailing lower block update (rank 16)
        for (int q = tid; q < rest * rest; q += 256) {
            const int rr = c0 + 16 + q % rest, cc = c0 + 16 + q / rest;
            if (rr < cc) continue;
            double sv = 0.0;
#pragma unroll
            for (int i = 0; i < 16; i++) sv += T[rr][c0 + i] * T[cc][c0 + i];
            T[rr][cc] -= sv;
        }
        __syncthreads();
    }
    // ---- X = L^-1 (lower): X(i,j), i > j, kept at T[j][i]; X(i,i) in xd[i] ----
    // diagonal 16x16 blocks: wave w inverts block w, lane c < 16 one column
    {
        const int c0 = wave * 16, c = lane;
        if (c < 16) {
            double x[16];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                if (i < c) {
                    x[i] = 0.0;
                    continue;
                }
                double sv = (i == c) ? 1.0 : 0.0;
#pragma unroll
                for (int k = 0; k < i; k++)
                    if (k >= c) sv -= T[c0 + i][c0 + k] * x[k];
                x[i] = sv / T[c0 + i][c0 + i];
            }
            xd[c0 + c] = x[c];
#pragma unroll
            for (int i = 0; i < 16; i++)
                if (i > c) T[c0 + c][c0 + i] = x[i];
        }
    }
    __syncthreads();
    // off-diagonal blocks by distance: X_IJ = -X_II sum_{K=J}^{I-1} L_IK X_KJ
    for (int dist = 1; dist < 4; dist++) {
        const int J = wave, I = wave + dist;
        const bool act = I < 4;
        const int a = lane & 15;
        if (act) {
            for (int t = 0; t < 4; t++) {
                const int bcol = (lane >> 4) + 4 * t;
                double sv = 0.0;
                for (int K = J; K < I; K++)
                    for (int k = 0; k < 16; k++) {
                        const int row = K * 16 + k, col = J * 16 + bcol;
                        double xv;
                        if (row > col) xv = T[col][row];
                        else xv = (row == col) ? xd[row] : 0.0;
                        sv += T[I * 16 + a][K * 16 + k] * xv;
                    }
                tmp[wave][a][bcol] = sv;
            }
        }
        __syncthreads();
        if (act) {
            for (int t = 0; t < 4; t++) {
                const int bcol = (lane >> 4) + 4 * t;
                double sv = 0.0;
                for (int k = 0; k <= a; k++) {
                    const int row = I * 16 + a, col = I * 16 + k;
                    const double xv = (row == col) ? xd[row] : T[col][row];
                    sv += xv * tmp[wave][k][bcol];
                }
                T[J * 16 + bcol][I * 16 + a] = -sv;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0 && bad && d.info[0] == 0) d.info[0] = k0 + bad;
    double *Dinv = d.Dinv + (size_t)step * TILE * TILE;
    for (int q = tid; q < TILE * TILE; q += 256) {
        int i = q & 63, j = q >> 6;
        Dinv[q] = i > j ? T[j][i] : (i == j ? xd[i] : 0.0);
        if (i < nb && j < nb && i >= j) Akk[i + (size_t)j * lda] = T[i][j];
    }
}

// column below the diagonal: A_ik = (A_ik - L_iw L_kw^T) inv(L_kk)^T for
// step < i <= step + band, window w = [max(0, i - band), step) (left-looking),
// then A_ii -= L_ik L_ik^T (right-looking for the diagonal tiles)
// Right-looking Cholesky steps for a few large matrices (which 1): the panel
// only scales the already updated column tiles, L_ik = A_ik Dinv_k^T, and a
// separate launch updates the whole trailing matrix, A_ij -= L_ik L_jk^T for
// step < j <= i, with every tile in parallel.  One matrix gets work spread
// over all tiles at every step instead of one long left-looking update per
// panel tile (the batched per-subdomain factorisations keep the left-looking
// steps, which have workgroups enough).
__global__ void __launch_bounds__(256) k_potrf_panel_rl(const DenseDesc *D, int step) {
    int bx, by;
    xcd_tile(bx, by);
    const DenseDesc &d = D[by];
    const int nt = d.ntm, n = d.m, lda = d.m;
    const int i = step + 1 + bx;
    if (i >= nt) return;
    const int k0 = step * TILE, i0 = i * TILE;
    const int nb = min(TILE, n - k0), mi = min(TILE, n - i0);
    __shared__ double lds[FUSED_LDS];
    d4 acc[4] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
    // C^T = Dinv_k C^T on the transposed tile, i.e. C = C Dinv_k^T
    strip_finish<true>(acc, lds, d.S + i0 + (size_t)k0 * lda, lda, nb, mi, d.Dinv + (size_t)step * TILE * TILE, 1.0);
}

__global__ void __launch_bounds__(256) k_potrf_trail(const DenseDesc *D, int step) {
    int bx, by;
    xcd_tile(bx, by);
    const DenseDesc &d = D[by];
    const int q = bx;
    int I = (int)((sqrtf(8.0f * q + 1.0f) - 1.0f) * 0.5f);
    while ((I + 1) * (I + 2) / 2 <= q) I++;
    while (I * (I + 1) / 2 > q) I--;
    const int J = q - I * (I + 1) / 2;
    const int i = step + 1 + I, j = step + 1 + J;
    if (i >= d.ntm) return;
    const int n = d.m, lda = d.m, k0 = step * TILE, i0 = i * TILE, j0 = j * TILE;
    const int nb = min(TILE, n - k0), mi = min(TILE, n - i0), mj = min(TILE, n - j0);
    tile_gemm<false, true>(d.S + i0 + (size_t)j0 * lda, lda, d.S + i0 + (size_t)k0 * lda, lda,
                           d.S + j0 + (size_t)k0 * lda, lda, mi, mj, nb, -1.0, 1.0);
}

__global__ void __launch_bounds__(256) k_potrf_panel(const DenseDesc *D, int which, int step) {
    int bx, by;
    xcd_tile(bx, by);
    const DenseDesc &d = D[by];
    double *A, *B;
    int lda, n, band, nt, ncols, ntc;
    pick(d, which, A, lda, n, band, nt, B, ncols, ntc);
    const int i = step + 1 + bx;
    if (i >= nt || i > step + band) return;
    const int k0 = step * TILE, i0 = i * TILE;
    const int nb = min(TILE, n - k0), mi = min(TILE, n - i0);
    const int j0 = max(0, i - band) * TILE;
    double *C = A + i0 + (size_t)k0 * lda;
    // transposed: T = L_kw L_iw^T (nb x mi), C^T = Dinv_k (C^T - T), i.e. C = (C - L_iw L_kw^T) Dinv_k^T
    __shared__ double lds[FUSED_LDS];
    d4 acc[4] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
    if (k0 > j0)
        strip_gemm<false, true>(acc, lds, A + k0 + (size_t)j0 * lda, lda, A + i0 + (size_t)j0 * lda, lda, nb, mi,
                                k0 - j0);
    d4 o[4];
    strip_finish<true, true>(acc, lds, C, lda, nb, mi, d.Dinv + (size_t)step * TILE * TILE, 1.0, o);
    // right-looking update of the diagonal tile of row block i by this column
    diag_update(o, lds, A + i0 + (size_t)i0 * lda, lda, mi);
}

// ---- left-looking forward substitution X = L^-1 B, block row i ----
// X_iJ = inv(L_ii) (B_iJ - L_iw X_wJ), window w = [max(first nonzero tile of
// column tile J, i - band), i)
__global__ void __launch_bounds__(256, 3) k_trsm_row(const DenseDesc *D, int which, int step) {
    int bx, by;
    xcd_tile(bx, by);
    const DenseDesc &d = D[by];
    double *A, *B;
    int lda, n, band, nt, ncols, ntc;
    pick(d, which, A, lda, n, band, nt, B, ncols, ntc);
    const int J0 = 2 * bx, i = step;
    if (J0 >= ntc || i >= nt) return;
    const int i0 = i * TILE, c0 = J0 * TILE;
    const int mi = min(TILE, n - i0), N = min(2 * TILE, ncols - c0);
    const double *Di = d.Dinv + (size_t)i * TILE * TILE;
    bool valid[2];
    int jst[2], jmin = n;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int J = J0 + t;
        valid[t] = J < ntc;
        const int cs = which == 0 ? (valid[t] ? d.colstart[J] : i + 1) : J;
        if (cs > i) valid[t] = false;
        if (valid[t] && which == 1 && J == i) {
            // W = L^-1 I: the right-hand side is the identity, so the diagonal tile is inv(L_ii)
            const int nj = min(TILE, ncols - J * TILE);
            double *Cd = B + i0 + (size_t)J * TILE * n;
            for (int q = threadIdx.x; q < TILE * TILE; q += 256) {
                const int r = q & 63, c = q >> 6;
                if (r < mi && c < nj) Cd[r + (size_t)c * n] = Di[q];
            }
            valid[t] = false;
        }
        jst[t] = max(cs, i - band) * TILE;
        if (valid[t]) jmin = min(jmin, jst[t]);
    }
    if (!valid[0] && !valid[1]) return;  // uniform over the workgroup
    __shared__ double lds[FUSED2_LDS];
    d4 acc[2][4];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = d4{0.0, 0.0, 0.0, 0.0};
    double *C = B + i0 + (size_t)c0 * n;
    if (i0 > jmin) {
        const int kfirst[2] = {valid[0] ? jst[0] - jmin : i0 - jmin, valid[1] ? jst[1] - jmin : i0 - jmin};
        strip2_gemm(acc, lds, A + i0 + (size_t)jmin * lda, lda, B + jmin + (size_t)c0 * n, n, mi, N, i0 - jmin,
                    kfirst);
    }
    // which 1: the identity has no entries off the diagonal tile (beta 0)
    strip2_finish(acc, lds, C, n, mi, N, Di, which == 1 ? 0.0 : 1.0, valid);
}

// one column tile per workgroup (few subdomains: twice the workgroups of the pairs)
__global__ void __launch_bounds__(256, 4) k_trsm_row1(const DenseDesc *D, int which, int step) {
    int bx, by;
    xcd_tile(bx, by);
    const DenseDesc &d = D[by];
    double *A, *B;
    int lda, n, band, nt, ncols, ntc;
    pick(d, which, A, lda, n, band, nt, B, ncols, ntc);
    const int J = bx, i = step;
    if (J >= ntc || i >= nt) return;
    const int cs = which == 0 ? d.colstart[J] : J;
    if (cs > i) return;
    const int i0 = i * TILE, c0 = J * TILE;
    const int mi = min(TILE, n - i0), nj = min(TILE, ncols - c0);
    const int j0 = max(cs, i - band) * TILE;
    double *C = B + i0 + (size_t)c0 * n;
    if (which == 1 && J == i) {
        // W = L^-1 I: the right-hand side is the identity, so the diagonal tile is inv(L_ii)
        const double *Di = d.Dinv + (size_t)i * TILE * TILE;
        for (int q = threadIdx.x; q < TILE * TILE; q += 256) {
            const int r = q & 63, c = q >> 6;
            if (r < mi && c < nj) C[r + (size_t)c * n] = Di[q];
        }
        return;
    }
    __shared__ double lds[FUSED_LDS];
    d4 acc[4] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
    if (i0 > j0)
        strip_gemm<false, false>(acc, lds, A + i0 + (size_t)j0 * lda, lda, B + j0 + (size_t)c0 * n, n, mi, nj,
                                 i0 - j0);
    // which 1: the identity has no entries off the diagonal tile (beta 0)
    strip_finish<false>(acc, lds, C, n, mi, nj, d.Dinv + (size_t)i * TILE * TILE, which == 1 ? 0.0 : 1.0);
}

// lower tiles of  S = -Y^T Y  (which 0)  or  S = W^T W  (which 1)
__global__ void __launch_bounds__(256) k_syrk(const DenseDesc *D, int which) {
    int bx, by;
    xcd_tile(bx, by);
    const DenseDesc &d = D[by];
    const int q = bx;
    int I = (int)((sqrtf(8.0f * q + 1.0f) - 1.0f) * 0.5f);
    while ((I + 1) * (I + 2) / 2 <= q) I++;
    while (I * (I + 1) / 2 > q) I--;
    const int J = q - I * (I + 1) / 2;
    if (I >= d.ntm) return;
    const int m = d.m, i0 = I * TILE, j0 = J * TILE;
    const int mi = min(TILE, m - i0), mj = min(TILE, m - j0);
    double *C = d.S + i0 + (size_t)j0 * m;
    if (which == 0) {
        // S = -Y^T Y (K_bb is added by k_add_kbb); a tile with no coupled
        // interior rows (K = 0) is written as zeros
        const int ks = min(max(d.colstart[I], d.colstart[J]) * TILE, d.ni);
        const double *Y = d.Kib;
        tile_gemm<true, false>(C, m, Y + ks + (size_t)i0 * d.ni, d.ni, Y + ks + (size_t)j0 * d.ni, d.ni, mi, mj,
                               d.ni - ks, -1.0, 0.0);
    } else {
        int ks = i0;  // W is lower triangular: rows < max(I,J)*64 of these columns vanish
        tile_gemm<true, false>(C, m, d.W + ks + (size_t)i0 * m, m, d.W + ks + (size_t)j0 * m, m, mi, mj, m - ks,
                               1.0, 0.0);
    }
}

// ---- W = L^-1 of the S factor (which 1) by recursive doubling over tiles ----
// L is in the lower triangle of S and the inverses of its diagonal tiles in
// Dinv (lower triangular, zeros above their diagonal).  The diagonal tiles of
// W are those inverses; level t (1, 2, 4, ... tiles) joins the pairs of
// t-tile blocks [b1, b1 + t), [b1 + t, b1 + 2t), b1 = 2tp:
//   [W11 0; W21 W22] = [L11 0; L21 L22]^-1,   W21 = -W22 (L21 W11),
// phase 0 forming T = L21 W11 (tile (i, j) over the K range where W11's
// column tile j is nonzero), stored transposed in the upper triangle of S,
// which nothing reads after the factorisation; phase 1 writing
// W21 = -W22 T (K range where W22's row tile i is nonzero).  Every tile of a
// level is independent, so each level is two launches over all subdomains
// instead of one forward-substitution step per tile row.
__global__ void __launch_bounds__(256) k_trtri_diag(const DenseDesc *D) {
    int bx, by;
    xcd_tile(bx, by);
    const DenseDesc &d = D[by];
    const int i = bx;
    if (i >= d.ntm) return;
    const int r0 = i * TILE, mi = min(TILE, d.m - r0);
    const double *Di = d.Dinv + (size_t)i * TILE * TILE;
    double *W = d.W + r0 + (size_t)r0 * d.m;
    for (int q = threadIdx.x; q < TILE * TILE; q += 256) {
        const int r = q & 63, c = q >> 6;
        if (r < mi && c < mi) W[r + (size_t)c * d.m] = Di[q];
    }
}

__global__ void __launch_bounds__(256) k_trtri_level(const DenseDesc *D, int t, int phase) {
    int bx, by;
    xcd_tile(bx, by);
    const DenseDesc &d = D[by];
    const int nt = d.ntm, m = d.m, tt = t * t;
    const int p = bx / tt, i = (bx % tt) / t, j = bx % t;  // pair, tile row in block 2, tile column in block 1
    const int b1 = 2 * t * p, b2 = b1 + t;
    if (b2 + i >= nt) return;
    const int r2 = (b2 + i) * TILE, c1 = (b1 + j) * TILE, k2 = b2 * TILE;
    const int mi = min(TILE, m - r2);
    if (phase == 0) {
        // T(r2.., c1..) = L(r2.., c1..k2) W(c1..k2, c1..), stored as S[c1 + x, r2 + y] = T(y, x)
        tile_gemm<true, true>(d.S + c1 + (size_t)r2 * m, m, d.W + c1 + (size_t)c1 * m, m, d.S + r2 + (size_t)c1 * m, m,
                              TILE, mi, k2 - c1, 1.0, 0.0);
    } else {
        // W(r2.., c1..) = -W(r2.., k2..r2+mi) T(k2..r2+mi, c1..)
        tile_gemm<false, true>(d.W + r2 + (size_t)c1 * m, m, d.W + r2 + (size_t)k2 * m, m, d.S + c1 + (size_t)k2 * m, m,
                               mi, TILE, r2 + mi - k2, -1.0, 0.0);
    }
}

__global__ void k_fill(double *p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// d[i] = s[i] (16-bit indices widened to int), eight per thread and step
__global__ void k_widen_u16(const uint16_t *s, int *d, int64_t n) {
    const int64_t n8 = n / 8;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n8; i += (int64_t)gridDim.x * blockDim.x) {
        const uint4 v = reinterpret_cast<const uint4 *>(s)[i];
        reinterpret_cast<int4 *>(d)[2 * i] = make_int4(v.x & 0xffff, v.x >> 16, v.y & 0xffff, v.y >> 16);
        reinterpret_cast<int4 *>(d)[2 * i + 1] = make_int4(v.z & 0xffff, v.z >> 16, v.w & 0xffff, v.w >> 16);
    }
    for (int64_t i = 8 * n8 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}


// d[off[s] + k] = img[pos[s] + k] for k < off[s + 1] - off[s]: the 16-bit
// column indices of subdomain s, at pos[s] in the image of the pinned blocks,
// widened into their place in the 32-bit CSR array
// (and the same for the K values, doubles, unconverted)
template <class S, class D>
__global__ void k_widen_runs(const S *img, const int64_t *pos, const int64_t *off, D *d) {
    const int s = blockIdx.y;
    const int64_t n = off[s + 1] - off[s];
    const S *src = img + pos[s];
    D *dst = d + off[s];
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = (D)src[k];
}

// K values handed over as the lower triangle (half[s] != 0: entries with
// column <= row, rows with sorted columns, so a prefix of each row).
// low[rp_off[s] + r] = offset of row r's lower entries in the subdomain's
// run: per row an upper bound of r among its columns, then a scan over the
// rows (one workgroup per subdomain, 256 rows per pass).
__global__ void __launch_bounds__(256) k_lower_offsets(const int *rowptr, const int64_t *rp_off, const int *colidx,
                                                       const int64_t *nz_off, const int *half, int *low) {
    const int s = blockIdx.x, tid = threadIdx.x;
    if (!half[s]) return;
    const int n = (int)(rp_off[s + 1] - rp_off[s]) - 1;
    const int *rp = rowptr + rp_off[s];
    const int *ci = colidx + nz_off[s];
    __shared__ int sh[256];
    int carry = 0;
    for (int base = 0; base < n; base += 256) {
        const int r = base + tid;
        int cnt = 0;
        if (r < n) {
            int lo = rp[r], hi = rp[r + 1];  // first column > r
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ci[mid] <= r) lo = mid + 1;
                else hi = mid;
            }
            cnt = lo - rp[r];
        }
        sh[tid] = cnt;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int t = tid >= o ? sh[tid - o] : 0;
            __syncthreads();
            sh[tid] += t;
            __syncthreads();
        }
        if (r < n) low[rp_off[s] + r] = carry + sh[tid] - cnt;
        carry += sh[255];
        __syncthreads();
    }
}

// vals[nz_off[s] + q] for every CSR entry q of subdomain s from its run at
// pos[s] of the value image: the run itself (half[s] == 0), or the lower
// entry (q, or the mirror of q found in row col(q)) through low (half[s]).
// Workgroup (g, s): rows 64g..64g+63, 16 lanes per row.  A mirror missing
// from the structure sets *err.
__global__ void __launch_bounds__(256) k_expand_sym(const double *img, const int64_t *pos, const int64_t *nz_off,
                                                    const int *rowptr, const int64_t *rp_off, const int *colidx,
                                                    const int *half, const int *low, double *vals, int *err) {
    const int s = blockIdx.y, g = blockIdx.x;
    const int n = (int)(rp_off[s + 1] - rp_off[s]) - 1;
    const int *rp = rowptr + rp_off[s];
    const int *ci = colidx + nz_off[s];
    const int *lw = low + rp_off[s];
    const double *src = img + pos[s];
    double *dst = vals + nz_off[s];
    const bool h = half[s] != 0;
    const int grp = threadIdx.x >> 4, l = threadIdx.x & 15;
    for (int k = 0; k < 4; k++) {
        const int r = g * 64 + 16 * k + grp;
        if (r >= n) break;
        for (int q = rp[r] + l; q < rp[r + 1]; q += 16) {
            if (!h) {
                dst[q] = src[q];
                continue;
            }
            const int c = ci[q];
            if (c <= r) {
                dst[q] = src[lw[r] + (q - rp[r])];
                continue;
            }
            int lo = rp[c], hi = rp[c + 1];  // r among row c's columns
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ci[mid] < r) lo = mid + 1;
                else hi = mid;
            }
            if (lo < rp[c + 1] && ci[lo] == r) {
                dst[q] = src[lw[c] + (lo - rp[c])];
            } else {
                dst[q] = 0.0;
                atomicOr(err, 1);
            }
        }
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

// S += rho Q Q^T on the lower triangle, one 64 x 64 tile per workgroup (the
// lower tiles in the order of k_syrk): the tile's rows and columns of Q (up to
// 8 kernel vectors) staged in LDS, so S is read and written once, coalesced
// along columns; kernel vectors beyond 8 are read from Q directly
__global__ void __launch_bounds__(256) k_kernel_reg(const DenseDesc *D, const double *Q, const int64_t *q_off,
                                                    const int *kdim, const double *rho) {
    constexpr int KR = 8;
    __shared__ double qr[KR][TILE], qc[KR][TILE];
    int bx, by;
    xcd_tile(bx, by);
    const DenseDesc &d = D[by];
    int I = (int)((sqrtf(8.0f * bx + 1.0f) - 1.0f) * 0.5f);
    while ((I + 1) * (I + 2) / 2 <= bx) I++;
    while (I * (I + 1) / 2 > bx) I--;
    const int J = bx - I * (I + 1) / 2;
    if (I >= d.ntm) return;
    const int m = d.m, k = kdim[by], i0 = I * TILE, j0 = J * TILE;
    const double r = rho[by];
    const double *q = Q + q_off[by];
    for (int t = threadIdx.x; t < KR * TILE; t += 256) {
        const int a = t / TILE, x = t % TILE;
        qr[a][x] = (a < k && i0 + x < m) ? q[i0 + x + (size_t)a * m] : 0.0;
        qc[a][x] = (a < k && j0 + x < m) ? q[j0 + x + (size_t)a * m] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < TILE * TILE; e += 256) {
        const int x = e & (TILE - 1), y = e / TILE, i = i0 + x, j = j0 + y;
        if (i >= m || j >= m || i < j) continue;
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < KR; a++) s += qr[a][x] * qc[a][y];
        for (int a = KR; a < k; a++) s += q[i + (size_t)a * m] * q[j + (size_t)a * m];
        d.S[i + (size_t)j * m] += r * s;
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
    // Y(:, j) is zero above its column tile's first coupled row (not stored)
    for (int r = d.colstart[j / TILE] * TILE + lane; r < d.ni; r += 64) s += y[r] * yf[r];
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

// nrm[b] = max_c sum_k |S(c, k)|, the infinity norm of the symmetric S held in
// its lower triangle (column-major, ld m): it bounds every partial sum of S x
// by nrm max|x| (the dual-operator SYMV scales its fixed-point accumulators
// with it).  Row c is its part left of the diagonal (S(c, k), k <= c: thread c
// walks the row, consecutive threads take consecutive rows, so the loads
// coalesce) plus the part below the diagonal (column c, read by one wave per
// column).  Fixed summation order; the max over the rows is exact, nrm[b]
// starts at zero.
__global__ void __launch_bounds__(256) k_rownorm(const DenseDesc *D, double *nrm) {
    const DenseDesc &d = D[blockIdx.y];
    const int m = d.m;
    const int c0 = blockIdx.x * 256;
    if (c0 >= m) return;
    __shared__ double below[256], wm[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double *S = d.S;
    for (int q = wave; q < 256; q += 4) {
        const int c = c0 + q;
        double v = 0.0;
        if (c < m)
            for (int i = c + 1 + lane; i < m; i += 64) v += fabs(S[i + (size_t)c * m]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) below[q] = v;
    }
    __syncthreads();
    const int c = c0 + tid;
    double r = 0.0;
    if (c < m) {
        for (int k = 0; k <= c; k++) r += fabs(S[c + (size_t)k * m]);
        r += below[tid];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o, 64));
    if (lane == 0) wm[wave] = r;
    __syncthreads();
    if (tid == 0) {
        const double b = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
        // non-negative doubles order like their bit patterns
        atomicMax((unsigned long long *)&nrm[blockIdx.y], (unsigned long long)__double_as_longlong(b));
    }
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
    // the VEC rows of one 16-byte group of the column, written at once: the
    // lanes of a wave write consecutive groups (1 KB contiguous)
#pragma unroll 2
    for (int g = 0; g < R / VEC; g++) {
        union {
            T v[VEC];
            uint4 q;
        } grp;
#pragma unroll
        for (int u = 0; u < VEC; u++) {
            const int row = r0 + g * VEC + u;
            double v = 0.0;
            if (row < m && row <= c) {
                v = d.S[c + (size_t)row * m];
                if (row == c) v *= 0.5;
            }
            v *= inv;
            if constexpr (sizeof(T) == 8) grp.v[u] = (T)v;
            else if constexpr (sizeof(T) == 4) grp.v[u] = (T)(float)v;
            else grp.v[u] = to_bf16((float)v);
        }
        *reinterpret_cast<uint4 *>(base + strip_index(m, R, VEC, t, g * VEC, c)) = grp.q;
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
        dst[e] = (c + dd < d.ni) ? d.Kii[(c + dd) + (size_t)c * d.ldk] : 0.0;
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


// Dense inputs of the assembly straight from the CSR rows, zeros included, in
// one coalesced pass (K is symmetric, so row perm[c] of K holds permuted
// column c): for interior column c the lower band tiles of K_ii, for boundary
// column j K_ib(:, j) from its first coupled row, and the permuted f as column
// m of [K_ib | f].  Every entry a later kernel reads is written here or by the
// SYRK (S), so the workspace needs no zero fill.
// One wave per column (4 columns per workgroup), each wave assembling its column
// in its own LDS window of MAT_W rows; only wave-level synchronisation, so the
// waves of a workgroup proceed independently.
constexpr int MAT_W = 1024;

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(256) k_materialize(const DenseDesc *D, const int *rowptr, const int64_t *rp_off,
                                                     const int *colidx, const double *vals, const int64_t *nz_off,
                                                     const int *iperm, const int *perm, const int64_t *perm_off,
                                                     const double *f, const int *nloc) {
    __shared__ double win[4][MAT_W];
    const DenseDesc &d = D[blockIdx.y];
    const int b = blockIdx.y, n = nloc[b], ni = d.ni, m = d.m;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = blockIdx.x * 4 + wave;
    if (c > n) return;  // columns 0..n-1 of K plus n = f; no workgroup barrier follows
    const int *rp = rowptr + rp_off[b];
    const int *ci = colidx + nz_off[b];
    const double *kv = vals + nz_off[b];
    const int *ip = iperm + perm_off[b];
    const int *pm = perm + perm_off[b];
    if (c == n) {  // f column, permuted
        double *dst = d.Kib + (size_t)m * ni;
        for (int r = lane; r < ni; r += 64) dst[r] = f[perm_off[b] + pm[r]];
        return;
    }
    // rows [lo, hi) of the permuted column c
    int lo, hi;
    if (c < ni) {
        lo = (c / TILE) * TILE;
        hi = min(ni, lo + (d.band_ii + 1) * TILE);
    } else {
        // K_ib(:, j) from the column tile's first coupled interior row: the
        // rows above it are zero in K_ib and in Y = L^-1 K_ib, and no kernel
        // reads them (the substitution, the SYRK and h_f start there).  K_bb is
        // not written here: S = -Y^T Y is written whole by the SYRK and K_bb's
        // entries are added to it afterwards (k_add_kbb), so the zeros of S
        // cross HBM once instead of three times.
        lo = d.colstart[(c - ni) / TILE] * TILE;
        hi = ni;
    }
    const int l = pm[c];
    const int q0 = rp[l], q1 = rp[l + 1];
    double *cb = win[wave];
    for (int w0 = lo; w0 < hi; w0 += MAT_W) {
        const int w1 = min(hi, w0 + MAT_W);
        for (int r = w0 + lane; r < w1; r += 64) cb[r - w0] = 0.0;
        wave_lds_sync();
        for (int q = q0 + lane; q < q1; q += 64) {
            const int pr = ip[ci[q]];
            if (pr >= w0 && pr < w1) cb[pr - w0] = kv[q];
        }
        wave_lds_sync();
        double *dst = c < ni ? d.Kii + (size_t)c * d.ldk : d.Kib + (size_t)(c - ni) * ni;
        for (int r = w0 + lane; r < w1; r += 64) dst[r] = cb[r - w0];
        wave_lds_sync();  // the window is read before the next one clears it
    }
}

// S += K_bb on the lower tiles (rows from the diagonal tile of each column
// down, the region the SYRK writes), after S = -Y^T Y: one wave per boundary
// column walks its CSR row; the entries of a column are distinct, so no two
// lanes add to one element.
__global__ void __launch_bounds__(256) k_add_kbb(const DenseDesc *D, const int *rowptr, const int64_t *rp_off,
                                                 const int *colidx, const double *vals, const int64_t *nz_off,
                                                 const int *iperm, const int *perm, const int64_t *perm_off) {
    const DenseDesc &d = D[blockIdx.y];
    const int b = blockIdx.y, ni = d.ni, m = d.m;
    const int lane = threadIdx.x & 63, j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= m) return;
    const int *rp = rowptr + rp_off[b];
    const int *ci = colidx + nz_off[b];
    const double *kv = vals + nz_off[b];
    const int *ip = iperm + perm_off[b];
    const int l = perm[perm_off[b] + ni + j];
    const int s0 = ni + (j / TILE) * TILE;  // first S row of the column's lower tiles, permuted
    double *col = d.S + (size_t)j * m - ni;
    for (int q = rp[l] + lane; q < rp[l + 1]; q += 64) {
        const int pr = ip[ci[q]];
        if (pr >= s0) col[pr] += kv[q];
    }
}

// ---- hybrid FETI: F_i B0_i^T and B0_i F_i B0_i^T (S holds F_i) ----
// Slot q of a subdomain is one B0 row touching it, with entries (boundary
// index j_e, coefficient c_e): FB0[j + q m] = sum_e c_e F(j, j_e).  With
// `sym` S holds F in both triangles (column j_e read contiguously), else in
// its lower triangle only (one entry per slot: corners).
__global__ void k_b0_columns(const DenseDesc *D, const int *nq, const int64_t *qoff, const int64_t *qe_ptr,
                             const int *qe_j, const double *qe_c, const int64_t *fb0off, double *FB0, int sym) {
    const DenseDesc &d = D[blockIdx.y];
    const int b = blockIdx.y, m = d.m, n = nq[b];
    double *fb = FB0 + fb0off[b];
    const int64_t tot = (int64_t)m * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e % m), q = (int)(e / m);
        const int64_t slot = qoff[b] + q;
        double v = 0.0;
        for (int64_t k = qe_ptr[slot]; k < qe_ptr[slot + 1]; k++) {
            const int c = qe_j[k];
            v += qe_c[k] * ((sym || j >= c) ? d.S[j + (size_t)c * m] : d.S[c + (size_t)j * m]);
        }
        fb[e] = v;
    }
}

// Fcc[q1 + q2 nq] = sum over the entries e of slot q1 of c_e FB0[j_e + q2 m]
__global__ void k_b0_block(const DenseDesc *D, const int *nq, const int64_t *qoff, const int64_t *qe_ptr,
                           const int *qe_j, const double *qe_c, const int64_t *fb0off, const int64_t *fccoff,
                           const double *FB0, double *Fcc) {
    const DenseDesc &d = D[blockIdx.x];
    const int b = blockIdx.x, m = d.m, n = nq[b];
    const double *fb = FB0 + fb0off[b];
    double *fc = Fcc + fccoff[b];
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int q1 = e % n, q2 = e / n;
        const int64_t slot = qoff[b] + q1;
        double v = 0.0;
        for (int64_t k = qe_ptr[slot]; k < qe_ptr[slot + 1]; k++) v += qe_c[k] * fb[qe_j[k] + (int64_t)q2 * m];
        fc[e] = v;
    }
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

__global__ void k_probe_mfma4(const double *a, const double *b, double *out) {
    const int lane = threadIdx.x;
    out[lane] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[lane], b[lane], 0.0, 0, 0, 0);
}

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_probe_gemm(int M, int N, int K, const double *A, int lda, const double *B,
                                                    int ldb, double *C, int ldc, double alpha, double beta) {
    tile_gemm<TA, TB>(C, ldc, A, lda, B, ldb, M, N, K, alpha, beta);
}

// FP64 MFMA issue rate: CH independent accumulator chains per wave, `iter`
// steps each, register operands only (one store per lane at the end).  SHAPE 0:
// v_mfma_f64_16x16x4f64 (2048 flops), 1: v_mfma_f64_4x4x4f64 with 4 blocks (512),
// both on one pair of operand registers.  SHAPE 2 (4x4x4) and 3 (16x16x4): the
// operand grid of a register-blocked tile, CH A and CH B registers and CH x CH
// accumulators, every MFMA on another operand pair.
template <int SHAPE, int CH>
__global__ void __launch_bounds__(256) k_mfma_rate(double *out, int iter) {
    const double a = 1.0 + threadIdx.x * 1e-9, b = 1.0 - threadIdx.x * 1e-9;
    double s = 0.0;
    if constexpr (SHAPE == 2) {
        double av[CH], bv[CH], acc[CH][CH];
#pragma unroll
        for (int r = 0; r < CH; r++) {
            av[r] = a + r * 1e-7;
            bv[r] = b - r * 1e-7;
#pragma unroll
            for (int c = 0; c < CH; c++) acc[r][c] = 0.0;
        }
        for (int i = 0; i < iter; i++)
#pragma unroll
            for (int r = 0; r < CH; r++)
#pragma unroll
                for (int c = 0; c < CH; c++) acc[r][c] = __builtin_amdgcn_mfma_f64_4x4x4f64(av[r], bv[c], acc[r][c], 0, 0, 0);
#pragma unroll
        for (int r = 0; r < CH; r++)
#pragma unroll
            for (int c = 0; c < CH; c++) s += acc[r][c];
    } else if constexpr (SHAPE == 3) {
        double av[CH], bv[CH];
        d4 acc[CH][CH];
#pragma unroll
        for (int r = 0; r < CH; r++) {
            av[r] = a + r * 1e-7;
            bv[r] = b - r * 1e-7;
#pragma unroll
            for (int c = 0; c < CH; c++) acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};
        }
        for (int i = 0; i < iter; i++)
#pragma unroll
            for (int r = 0; r < CH; r++)
#pragma unroll
                for (int c = 0; c < CH; c++)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r], bv[c], acc[r][c], 0, 0, 0);
#pragma unroll
        for (int r = 0; r < CH; r++)
#pragma unroll
            for (int c = 0; c < CH; c++) s += (acc[r][c][0] + acc[r][c][1]) + (acc[r][c][2] + acc[r][c][3]);
    } else if constexpr (SHAPE == 0) {
        d4 acc[CH];
#pragma unroll
        for (int c = 0; c < CH; c++) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < iter; i++)
#pragma unroll
            for (int c = 0; c < CH; c++) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < CH; c++) s += (acc[c][0] + acc[c][1]) + (acc[c][2] + acc[c][3]);
    } else {
        double acc[CH];
#pragma unroll
        for (int c = 0; c < CH; c++) acc[c] = 0.0;
        for (int i = 0; i < iter; i++)
#pragma unroll
            for (int c = 0; c < CH; c++) acc[c] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc[c], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < CH; c++) s += acc[c];
    }
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

double mfma_rate_flops(int shape) { return (shape == 0 || shape == 3) ? 2048.0 : 512.0; }

void launch_mfma_rate(int shape, int chains, int blocks, int iter, int lds, double *out, hipStream_t st) {
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), lds, st, out, iter); };
    if (shape == 2) {
        if (chains == 2) go(k_mfma_rate<2, 2>);
        else if (chains == 4) go(k_mfma_rate<2, 4>);
        else go(k_mfma_rate<2, 8>);
    } else if (shape == 3) {
        if (chains == 2) go(k_mfma_rate<3, 2>);
        else go(k_mfma_rate<3, 4>);
    } else if (shape == 0) {
        if (chains == 4) go(k_mfma_rate<0, 4>);
        else if (chains == 8) go(k_mfma_rate<0, 8>);
        else if (chains == 16) go(k_mfma_rate<0, 16>);
        else go(k_mfma_rate<0, 32>);
    } else {
        if (chains == 4) go(k_mfma_rate<1, 4>);
        else if (chains == 8) go(k_mfma_rate<1, 8>);
        else if (chains == 16) go(k_mfma_rate<1, 16>);
        else go(k_mfma_rate<1, 32>);
    }
}

void launch_probe_mfma(const double *a, const double *b, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_probe_mfma, dim3(1), dim3(64), 0, st, a, b, out);
    hipLaunchKernelGGL(k_probe_mfma4, dim3(1), dim3(64), 0, st, a, b, out + 256);
}
void launch_probe_gemm(int ta, int tb, int M, int N, int K, const double *A, int lda, const double *B, int ldb,
                       double *C, int ldc, double alpha, double beta, hipStream_t st) {
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(1), dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc,
                                                  alpha, beta); };
    if (!ta && !tb) go(k_probe_gemm<false, false>);
    else if (!ta && tb) go(k_probe_gemm<false, true>);
    else if (ta && !tb) go(k_probe_gemm<true, false>);
    else go(k_probe_gemm<true, true>);
}

// ------------------------------------------------------------ launchers ----
static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

void launch_widen_runs(const uint16_t *img, const int64_t *pos, const int64_t *off, int nsub, int64_t max_n, int *d,
                       hipStream_t st) {
    if (nsub <= 0 || max_n <= 0) return;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_n + 1023) / 1024, 256));
    hipLaunchKernelGGL((k_widen_runs<uint16_t, int>), dim3(g, nsub), dim3(256), 0, st, img, pos, off, d);
}

void launch_copy_runs(const double *img, const int64_t *pos, const int64_t *off, int nsub, int64_t max_n, double *d,
                      hipStream_t st) {
    if (nsub <= 0 || max_n <= 0) return;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_n + 1023) / 1024, 256));
    hipLaunchKernelGGL((k_widen_runs<double, double>), dim3(g, nsub), dim3(256), 0, st, img, pos, off, d);
}

void launch_expand_sym(const double *img, const int64_t *pos, const int64_t *nz_off, const int *rowptr,
                       const int64_t *rp_off, const int *colidx, const int *half, int *low, int nsub, int max_n,
                       double *vals, int *err, hipStream_t st) {
    if (nsub <= 0 || max_n <= 0) return;
    hipLaunchKernelGGL(k_lower_offsets, dim3(nsub), dim3(256), 0, st, rowptr, rp_off, colidx, nz_off, half, low);
    hipLaunchKernelGGL(k_expand_sym, dim3((max_n + 63) / 64, nsub), dim3(256), 0, st, img, pos, nz_off, rowptr, rp_off,
                       colidx, half, low, vals, err);
}

void launch_widen_u16(const uint16_t *s, int *d, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n / 8 + 255) / 256, 8192));
    hipLaunchKernelGGL(k_widen_u16, dim3(g), dim3(256), 0, st, s, d, n);
}

void launch_fill(double *p, int64_t n, double v, hipStream_t st) {
    if (n <= 0) return;
    unsigned g = (unsigned)std::min<int64_t>(cdiv(n, 256), 65536);
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, st, p, n, v);
}

void launch_potrf_diag(const DenseDesc *desc, int nb, int which, int step, hipStream_t st) {
    hipLaunchKernelGGL(k_potrf_diag, dim3(nb), dim3(256), 0, st, desc, which, step);
}

void launch_potrf_panel(const DenseDesc *desc, int nb, int which, int step, int max_band, hipStream_t st) {
    if (max_band <= 0) return;
    hipLaunchKernelGGL(k_potrf_panel, dim3(max_band, nb), dim3(256), 0, st, desc, which, step);
}

void launch_potrf_rl(const DenseDesc *desc, int nb, int max_ntm, hipStream_t st) {
    for (int k = 0; k < max_ntm; k++) {
        launch_potrf_diag(desc, nb, 1, k, st);
        const int T = max_ntm - k - 1;
        if (T <= 0) break;
        hipLaunchKernelGGL(k_potrf_panel_rl, dim3(T, nb), dim3(256), 0, st, desc, k);
        hipLaunchKernelGGL(k_potrf_trail, dim3(T * (T + 1) / 2, nb), dim3(256), 0, st, desc, k);
    }
}

void launch_potrf_step(const DenseDesc *desc, int nb, int which, int step, int max_band, hipStream_t st) {
    launch_potrf_diag(desc, nb, which, step, st);
    launch_potrf_panel(desc, nb, which, step, max_band, st);
}

void launch_trsm_step(const DenseDesc *desc, int nb, int which, int step, int max_band, int max_ntc,
                      hipStream_t st) {
    (void)max_band;
    // pairs of column tiles when they still fill the GPU (256 CUs, 3 workgroups each)
    if ((int64_t)nb * ((max_ntc + 1) / 2) >= 768)
        hipLaunchKernelGGL(k_trsm_row, dim3((max_ntc + 1) / 2, nb), dim3(256), 0, st, desc, which, step);
    else
        hipLaunchKernelGGL(k_trsm_row1, dim3(max_ntc, nb), dim3(256), 0, st, desc, which, step);
}

void launch_trtri(const DenseDesc *desc, int nb, int max_ntm, hipStream_t st) {
    hipLaunchKernelGGL(k_trtri_diag, dim3(max_ntm, nb), dim3(256), 0, st, desc);
    for (int t = 1; t < max_ntm; t *= 2) {
        const unsigned g = (unsigned)(cdiv(max_ntm, 2 * t) * t * t);
        hipLaunchKernelGGL(k_trtri_level, dim3(g, nb), dim3(256), 0, st, desc, t, 0);
        hipLaunchKernelGGL(k_trtri_level, dim3(g, nb), dim3(256), 0, st, desc, t, 1);
    }
}

void launch_syrk(const DenseDesc *desc, int nb, int which, int max_ntm, hipStream_t st) {
    hipLaunchKernelGGL(k_syrk, dim3(max_ntm * (max_ntm + 1) / 2, nb), dim3(256), 0, st, desc, which);
}


void launch_kernel_reg(const DenseDesc *desc, int nb, const double *q, const int64_t *q_off, const int *kdim,
                       double *rho, int max_m, hipStream_t st) {
    hipLaunchKernelGGL(k_rho, dim3(nb), dim3(256), 0, st, desc, rho);
    const int nt = (max_m + TILE - 1) / TILE;
    hipLaunchKernelGGL(k_kernel_reg, dim3(nt * (nt + 1) / 2, nb), dim3(256), 0, st, desc, q, q_off, kdim, rho);
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

void launch_rownorm(const DenseDesc *desc, int nb, double *nrm, int max_m, hipStream_t st) {
    launch_fill(nrm, nb, 0.0, st);
    hipLaunchKernelGGL(k_rownorm, dim3(cdiv(max_m, 256), nb), dim3(256), 0, st, desc, nrm);
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

void launch_add_kbb(const DenseDesc *desc, int nb, const int *rowptr, const int64_t *rp_off, const int *colidx,
                    const double *vals, const int64_t *nz_off, const int *iperm, const int *perm,
                    const int64_t *perm_off, int max_m, hipStream_t st) {
    hipLaunchKernelGGL(k_add_kbb, dim3(cdiv(max_m, 4), nb), dim3(256), 0, st, desc, rowptr, rp_off, colidx, vals,
                       nz_off, iperm, perm, perm_off);
}

void launch_materialize(const DenseDesc *desc, int nb, const int *rowptr, const int64_t *rp_off, const int *colidx,
                        const double *vals, const int64_t *nz_off, const int *iperm, const int *perm,
                        const int64_t *perm_off, const double *f, const int *n, int max_n, hipStream_t st) {
    hipLaunchKernelGGL(k_materialize, dim3(cdiv(max_n + 1, 4), nb), dim3(256), 0, st, desc, rowptr, rp_off, colidx,
                       vals, nz_off, iperm, perm, perm_off, f, n);
}


void launch_extract_b0(const DenseDesc *desc, int nb, const int *nq, const int64_t *qoff, const int64_t *qe_ptr,
                       const int *qe_j, const double *qe_c, const int64_t *fb0off, const int64_t *fccoff, double *FB0,
                       double *Fcc, int max_m, int max_nq, bool sym, hipStream_t st) {
    if (sym) launch_symmetrize_desc(desc, nb, max_m, st);
    const unsigned g = std::max<unsigned>(1, std::min<unsigned>(cdiv((int64_t)max_m * max_nq, 256), 4096));
    hipLaunchKernelGGL(k_b0_columns, dim3(g, nb), dim3(256), 0, st, desc, nq, qoff, qe_ptr, qe_j, qe_c, fb0off, FB0,
                       sym ? 1 : 0);
    hipLaunchKernelGGL(k_b0_block, dim3(nb), dim3(256), 0, st, desc, nq, qoff, qe_ptr, qe_j, qe_c, fb0off, fccoff, FB0,
                       Fcc);
}

}  // namespace esp
