// Internal declarations shared by the host driver (feti.hip, grid.cpp) and the
// assembly kernels (dense.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/espreso_amd.h"

namespace esp {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define ESP_HIP(expr)                                                                                \
    do {                                                                                             \
        hipError_t _e = (expr);                                                                      \
        if (_e != hipSuccess)                                                                        \
            throw ::esp::Error(ESPRESO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define ESP_CHECK(cond, code, msg)                      \
    do {                                                \
        if (!(cond)) throw ::esp::Error((code), (msg)); \
    } while (0)

constexpr int TILE = 64;        // dense tile edge of the assembly kernels
constexpr int STRIP_F64 = 16;   // rows per packed strip, FP64 storage
constexpr int STRIP_F32 = 32;   // FP32 storage
constexpr int STRIP_B16 = 64;   // BF16 storage
// SYMV work-item sizes (matrix elements per workgroup): default and autotuner candidates
constexpr int64_t SYMV_DEFAULT_TARGET = 320 * 1024;
constexpr int64_t SYMV_TUNE_TARGETS[5] = {96 * 1024, 192 * 1024, 320 * 1024, 640 * 1024, 1280 * 1024};

__host__ __device__ inline int strip_rows(int prec) { return prec == 0 ? STRIP_F64 : (prec == 1 ? STRIP_F32 : STRIP_B16); }
// LDS of a SYMV workgroup (interface up to max_m DOFs): the 64-bit fixed-point
// row/column sums and x of the columns (plus R padding) in the accumulation
// type (FP64 for FP64 storage, FP32 otherwise)
__host__ __device__ inline int64_t symv_lds_bytes(int64_t max_m, int prec) {
    return 8 * max_m + (max_m + strip_rows(prec)) * (prec ? 4 : 8);
}

// Packed-strip layout of a symmetric m x m matrix with R rows per strip:
// strip t holds rows [tR, tR+R) x columns [tR, m) (R*(m-tR) values), zero
// below the diagonal, diagonal stored halved so that one pass applies both
// triangles.  Inside a strip the columns go in chunks of 64; a chunk of width
// w is stored [R/VEC][w][VEC] with VEC = 16 bytes / element size, so that the
// q-th 16-byte load of wave lane l (column l of the chunk) is contiguous with
// lane l+1's: every wave-wide load instruction reads one contiguous segment.
__host__ __device__ inline int64_t strip_offset(int64_t m, int64_t R, int64_t t) { return R * (t * m - R * t * (t - 1) / 2); }
__host__ __device__ inline int64_t strip_count(int64_t m, int64_t R) { return (m + R - 1) / R; }
__host__ __device__ inline int64_t strip_size(int64_t m, int64_t R) { return strip_offset(m, R, strip_count(m, R)); }
// position of element (row tR + r, column c >= tR) of strip t
__host__ __device__ inline int64_t strip_index(int64_t m, int64_t R, int64_t VEC, int64_t t, int64_t r, int64_t c) {
    const int64_t cl = c - t * R, k = cl >> 6, ncol = m - t * R;
    const int64_t w = (ncol - 64 * k) < 64 ? (ncol - 64 * k) : 64;
    return strip_offset(m, R, t) + 64 * k * R + ((r / VEC) * w + (cl & 63)) * VEC + r % VEC;
}

// Per-subdomain descriptor of the dense assembly workspace (device memory).
// Interior DOFs first (ascending local DOF), boundary DOFs after them.
struct DenseDesc {
    // K_ii (lower band, becomes L: K_ii = L L^T) in diagonal-band storage:
    // element (i, j) at Kii[i + j ldk] for -63 <= i - j <= ldk - 63, with
    // ldk + 1 = min((band_ii + 2) 64, ni + 64) stored rows per column and Kii
    // 63 elements past the start of the storage.  The tiles of the band, and every
    // window the kernels read, address like a column-major matrix of leading
    // dimension ldk; the storage is ni (ldk + 1) values instead of ni^2.
    double *Kii;
    double *Kib;          // ni x (m+1), ld ni: [K_ib | f_i]  ->  [Y | y_f] = L^-1 [K_ib | f_i]
    double *S;            // m x m, ld m (lower): K_bb - Y^T Y -> + rho Q Q^T -> L_s -> F
    double *W;            // m x m, ld m: L_s^-1 (lower tiles); may reuse Kib's storage
                          // (Y and y_f are no longer read once S_D and the band are out)
    double *Dinv;         // max(ntii, ntm) x 64 x 64: inverses of diagonal blocks
    const int *colstart;  // ntc: first tile row of a nonzero in each column tile of K_ib
    int *info;            // factorisation status (0 = ok, else 1-based failing column)
    int ni, m;
    int ldk;              // leading dimension of Kii (band storage, see above)
    int band_ii;          // block bandwidth of K_ii in tiles
    int ntii, ntm, ntc;   // tile counts: ceil(ni/64), ceil(m/64), ceil((m+1)/64)
};

// Launchers (dense.hip).  `desc` is a device array of nb descriptors.
// which = 0: operate on K_ii / [K_ib|f];  which = 1: on S / W.
void launch_fill(double *p, int64_t n, double v, hipStream_t st);
// d[i] = s[i] for i < n; s and d 16-byte aligned
void launch_widen_u16(const uint16_t *s, int *d, int64_t n, hipStream_t st);
// d[off[s] + k] = img[pos[s] + k], k < off[s + 1] - off[s] (max_n: the largest run)
void launch_widen_runs(const uint16_t *img, const int64_t *pos, const int64_t *off, int nsub, int64_t max_n, int *d,
                       hipStream_t st);
// K values from lower-triangle runs (half[s]) or full runs: low is scratch of
// rp_off[nsub] ints (per-row offsets), *err set when a mirror is missing
void launch_expand_sym(const double *img, const int64_t *pos, const int64_t *nz_off, const int *rowptr,
                       const int64_t *rp_off, const int *colidx, const int *half, int *low, int nsub, int max_n,
                       double *vals, int *err, hipStream_t st);
void launch_copy_runs(const double *img, const int64_t *pos, const int64_t *off, int nsub, int64_t max_n, double *d,
                      hipStream_t st);
void launch_potrf_step(const DenseDesc *desc, int nb, int which, int step, int max_band, hipStream_t st);
// the two halves of a step: factor and invert the diagonal tile, then the panel below it
void launch_potrf_diag(const DenseDesc *desc, int nb, int which, int step, hipStream_t st);
void launch_potrf_panel(const DenseDesc *desc, int nb, int which, int step, int max_band, hipStream_t st);
// whole right-looking Cholesky of S (which 1) for a few large matrices: per
// step the diagonal tile, the scaled column below it and the trailing update
void launch_potrf_rl(const DenseDesc *desc, int nb, int max_ntm, hipStream_t st);
void launch_trsm_step(const DenseDesc *desc, int nb, int which, int step, int max_band, int max_ntc,
                      hipStream_t st);
void launch_syrk(const DenseDesc *desc, int nb, int which, int max_ntm, hipStream_t st);
// W = L^-1 for the S factor (which 1: L in S's lower triangle, diagonal tile
// inverses in Dinv) by recursive doubling; uses S's upper triangle as scratch
void launch_trtri(const DenseDesc *desc, int nb, int max_ntm, hipStream_t st);
void launch_symmetrize_desc(const DenseDesc *desc, int nb, int max_m, hipStream_t st);
// CSR K_i (local DOFs) into the dense K_ii band and [K_ib|f], zeros included;
// iperm maps local -> internal index, perm internal -> local
void launch_add_kbb(const DenseDesc *desc, int nb, const int *rowptr, const int64_t *rp_off, const int *colidx,
                    const double *vals, const int64_t *nz_off, const int *iperm, const int *perm,
                    const int64_t *perm_off, int max_m, hipStream_t st);
void launch_materialize(const DenseDesc *desc, int nb, const int *rowptr, const int64_t *rp_off, const int *colidx,
                        const double *vals, const int64_t *nz_off, const int *iperm, const int *perm,
                        const int64_t *perm_off, const double *f, const int *n, int max_n, hipStream_t st);
// S += rho Q Q^T with rho = max diag S; Q (m x k, column-major) at q[q_off[b]]
void launch_kernel_reg(const DenseDesc *desc, int nb, const double *q, const int64_t *q_off,
                       const int *kdim, double *rho, int max_m, hipStream_t st);
// hf = fb - Y^T y_f  (fb, hf: m values at off[b])
void launch_hf(const DenseDesc *desc, int nb, const double *fb, double *hf, const int64_t *off, int max_m,
               hipStream_t st);
// pack the lower triangle of S into strips: prec 0 FP64, 1 FP32, 2 BF16 (scaled by 1/max|S|)
void launch_rownorm(const DenseDesc *desc, int nb, double *nrm, int max_m, hipStream_t st);
void launch_pack(const DenseDesc *desc, int nb, int prec, void *pool, const int64_t *off, float *scale,
                 int max_m, hipStream_t st);
// copy band of L (ni x (bw+1), column c holds L[c..c+bw][c]) and y_f out of the workspace
void launch_band_extract(const DenseDesc *desc, int nb, double *band, const int64_t *band_off, const int *bw,
                         double *yf, const int64_t *yf_off, int max_ni, hipStream_t st);
// unpack one strip-packed matrix into dense m x m (column-major, full)
void launch_unpack(const void *pool, int prec, int64_t off, int m, float scale, double *out, hipStream_t st);
// hybrid FETI: FB0 = F_i B0_i^T (m x nq, signed corner columns) and the
// corner block B0_i F_i B0_i^T (nq x nq), from F_i in the lower triangle of S
void launch_extract_b0(const DenseDesc *desc, int nb, const int *nq, const int64_t *qoff, const int64_t *qe_ptr,
                       const int *qe_j, const double *qe_c, const int64_t *fb0off, const int64_t *fccoff, double *FB0,
                       double *Fcc, int max_m, int max_nq, bool sym, hipStream_t st);

// diagnostics (dense.hip)
void launch_probe_mfma(const double *a, const double *b, double *out, hipStream_t st);
// FP64 MFMA issue-rate loop (shape 0: 16x16x4, 1: 4x4x4 with 4 blocks; chains 4 or 8);
// out holds blocks x 256 doubles
void launch_mfma_rate(int shape, int chains, int blocks, int iter, int lds, double *out, hipStream_t st);
double mfma_rate_flops(int shape);
void launch_probe_gemm(int ta, int tb, int M, int N, int K, const double *A, int lda, const double *B, int ldb,
                       double *C, int ldc, double alpha, double beta, hipStream_t st);

}  // namespace esp
