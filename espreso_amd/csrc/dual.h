// Device-resident description of the dual problem and launchers of the
// dual-space kernels (dual.hip).  Internal.
#pragma once

#include "internal.h"

namespace esp {

// One SYMV work item: a workgroup applies strips [t0, t1) of subdomain `sub`
// and writes its partial result (length m - t0*R) at part_off.
struct SymvItem {
    int sub;
    int t0, t1;
    int pad;
    int64_t part_off;
};

// Block SYMV work list (FP64 geometry, at most BLOCK_STRIPS strips per item)
// and its scratch: xb (nbnd x Kmax, right-hand sides interleaved) and the
// partials (plen x Kmax)
constexpr int BLOCK_STRIPS = 8;
struct SymvBlock {
    SymvItem *items = nullptr;
    int n_items = 0;
    int *item0 = nullptr, *nitems = nullptr;
    int64_t plen = 0;
    double *xb = nullptr, *ypart = nullptr;
};

struct DualDev {
    int nsub = 0, nlam = 0, ncoarse = 0, max_m = 0;
    int64_t nbnd = 0;                // sum of m_s
    // per subdomain
    int *m = nullptr;
    int64_t *boff = nullptr;         // offset into boundary arrays
    int *kdim = nullptr;
    int *coff = nullptr;             // coarse offset
    int64_t *rboff = nullptr;        // offset of R_b (m x k, column-major)
    int64_t *foff = nullptr;         // offset of packed F_i in its storage type
    int64_t *sdoff = nullptr;        // offset of packed S_D (FP64)
    float *fscale = nullptr;         // BF16 storage scale
    // infinity norms of F_i and S_D (FP64 values): the SYMV's fixed-point scale
    double *fnrm = nullptr, *sdnrm = nullptr;
    // boundary -> lambda (CSR over all boundary DOFs)
    int64_t *bl_ptr = nullptr;
    int *bl_lam = nullptr;
    double *bl_sign = nullptr;
    double *bl_sscale = nullptr;     // sign * scaling weight
    // lambda -> boundary entries (1 or 2 per lambda, slots 2l, 2l+1)
    int *le_cnt = nullptr;
    int64_t *le_b = nullptr;
    int *le_sub = nullptr;
    double *le_sign = nullptr;
    double *le_sscale = nullptr;
    // K_bb in boundary indices (lumped preconditioner)
    int64_t *kbb_ptr = nullptr, *kbb_col = nullptr;
    double *kbb_val = nullptr;
    // matrices
    double *Rb = nullptr;
    double *Fd = nullptr;            // packed F_i, FP64 (always kept)
    float *Ff = nullptr;             // packed F_i, FP32
    uint16_t *Fh = nullptr;          // packed F_i, BF16
    double *SD = nullptr;            // packed S_D, FP64 (sd_prec 0)
    float *SDf = nullptr;            // packed S_D, FP32 (sd_prec 1), low-precision geometry
    uint16_t *SDh = nullptr;         // packed S_D, BF16 (sd_prec 2), scaled by sdscale
    // offsets of the packed pools per storage precision (FP64, FP32, BF16 strip
    // geometry R = 16, 32, 64); poff[0] == foff, unused precisions null
    int64_t *poff[3] = {nullptr, nullptr, nullptr};
    float *sdscale = nullptr;
    int sd_prec = 0;
    double *GGinv = nullptr;
    // SYMV work lists per storage precision (F_i and S_D of one precision share them)
    SymvItem *items[3] = {nullptr, nullptr, nullptr};
    int n_items[3] = {0, 0, 0};
    int *item0[3] = {nullptr, nullptr, nullptr}, *nitems[3] = {nullptr, nullptr, nullptr};
    double *ypart = nullptr;
    double *yb = nullptr, *xb = nullptr;  // nbnd scratch
    double *cd = nullptr, *ce = nullptr;  // coarse scratch
    // coarse space: blocks (TFETI: subdomains, HTFETI: clusters); per subdomain
    // the basis Rt_s (m x ck[s]) restricted to the boundary and N_s (kdim x ck)
    // mapping block amplitudes to subdomain kernel amplitudes
    int nblk = 0;
    int *cblk = nullptr, *ck = nullptr, *blk_ptr = nullptr, *blk_sub = nullptr;
    int64_t *rtoff = nullptr, *dsub_off = nullptr, *noff = nullptr;
    double *Rt = nullptr, *N = nullptr, *dsub = nullptr;
    int64_t *asub = nullptr;          // offset of kernel amplitudes of subdomain s (sum of kdim)
};

// hybrid FETI clusters (B0 corners, F0, G0, Sa), device side
struct ClusterDev {
    int ncl = 0, max_n0 = 0, max_kt = 0, max_kn = 0;
    int *n0 = nullptr, *kt = nullptr, *kn = nullptr;  // B0 rows, sum of subdomain kernel dims, cluster kernel dim
    int *sfirst = nullptr, *scount = nullptr;        // consecutive subdomains of each cluster
    int64_t *rowoff = nullptr, *f0off = nullptr, *g0off = nullptr, *saoff = nullptr, *eoff = nullptr,
            *ncoff = nullptr;
    // entries of each cluster row (global rows rowoff[c] + r): global boundary
    // index and coefficient, CSR by row (g0 = B0 y_b)
    int64_t *re_ptr = nullptr, *re_b = nullptr;
    double *re_c = nullptr;
    double *F0inv = nullptr;   // n0 x n0 per cluster (full, symmetric)
    double *G0T = nullptr;     // n0 x kt per cluster: -(B0 R), i.e. G0^T
    double *H = nullptr;       // n0 x kt per cluster: F0^-1 G0^T
    double *Sainv = nullptr;   // kt x kt per cluster: (G0 F0^-1 G0^T + rho N N^T)^-1
    double *Nc = nullptr;      // kt x kn per cluster: orthonormal basis of ker G0^T (cluster rigid modes)
    // per subdomain: slots q = the B0 rows touching it (global row qrow, FB0
    // column q); the slot's entries (boundary index, coefficient) as a CSR over
    // all slots: corner rows have one entry per slot, kernel rows a face's worth
    int *nq = nullptr, *qrow = nullptr;
    int64_t *qoff = nullptr, *fb0off = nullptr, *fccoff = nullptr, *asub = nullptr;
    int64_t *qe_ptr = nullptr;
    int *qe_j = nullptr;
    double *qe_c = nullptr;
    double *FB0 = nullptr;     // m x nq per subdomain: F_i B0_i^T, column q = sum_e c_e F_i[:, j_e]
    float *FB0f = nullptr;     // FB0 in FP32 (FP32/BF16 storage of F_i only)
    double *Fcc = nullptr;     // nq x nq per subdomain: B0_i F_i B0_i^T
    double *lambda0 = nullptr, *alphaK = nullptr;
    double *rho_s = nullptr;   // per subdomain: max diag S_i (the scale of the kernel regularisation)
};

struct PrimalDev {
    int *ni = nullptr, *m = nullptr, *n = nullptr, *bw = nullptr, *kdim = nullptr, *coff = nullptr;
    int64_t *band_off = nullptr, *yf_off = nullptr, *perm_off = nullptr, *rp_off = nullptr, *nz_off = nullptr;
    int64_t *boff = nullptr, *work_off = nullptr, *r_off = nullptr;
    double *band = nullptr, *yf = nullptr, *vals = nullptr, *R = nullptr, *alpha = nullptr, *work = nullptr;
    int64_t *asub = nullptr;
    int *perm = nullptr, *iperm = nullptr, *rowptr = nullptr, *colidx = nullptr;
    int max_ni = 0;  // largest interior DOF count (host value: launch grids)
};

// y_b (D.yb) = A x_b for every subdomain; A = S_D (sd) or F_i in storage
// precision `prec` (0 FP64, 1 FP32, 2 BF16).  x is lambda-space and gathered
// with sign (or sign*scale when `scaled`) if `gather`, else boundary-space.
// ev0/ev1 (optional) are recorded around the SYMV kernel alone.
void launch_symv(const DualDev &D, bool sd, int prec, const double *x, int gather, bool scaled, hipStream_t st,
                 hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// nr right-hand sides at once on MFMA (1 <= nr <= 16; FP64 storage): x is
// lambda-space, nr vectors of n_lambda (gathered with sign, or sign*scale when
// `scaled`); yb receives nr boundary vectors of nbnd; stride nr in the scratch.
// ev0/ev1 bracket the SYMV kernel.
void launch_symv_mfma(const DualDev &D, const SymvBlock &B, bool sd, int nr, const double *x, bool scaled,
                      double *yb, hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void launch_lambda_gather(const DualDev &D, const double *yb, bool scaled, double *y, hipStream_t st);
void launch_boundary_gather(const DualDev &D, const double *x, bool scaled, double *xb, hipStream_t st);
void launch_kbb(const DualDev &D, const double *xb, double *yb, hipStream_t st);
// mode 0: y = x - G^T GGinv G x; mode 1: y = G^T GGinv x (x coarse); mode 2: y = GGinv G x (y coarse)
void launch_project(const DualDev &D, const double *x, double *y, int mode, hipStream_t st);
// out[asub[s] + a] = -R_s^T (B_s^T x) with the subdomain kernel (hybrid FETI e0)
void launch_subdomain_kernel_dots(const DualDev &D, const double *x, double *out, hipStream_t st);
// y_b += -K^+ B0^T lambda0 (+ R alphaK when addR) after solving the cluster problems
// lowp: apply -F_i B0^T lambda0 with FB0f (the inner iterations on the FP32/BF16 F_i)
void launch_cluster_correction(const DualDev &D, const ClusterDev &C, const double *e0s, double *yb, int addR,
                               hipStream_t st, bool lowp = false);
// FB0f = FB0 rounded to FP32 (for the inner iterations on FP32/BF16 storage)
void launch_fb0_round(const DualDev &D, const ClusterDev &C, hipStream_t st);
// the correction for nr cases of a block (case strides ystr, tstr/lstr, astr):
// t_v = F0^-1 B0 y_b,v into t (the lambda0 layout), then per case alpha and
// lambda0 (launch_cluster_alpha), then y_b,v += -F B0^T lambda0_v + R alphaK_v
void launch_cluster_t_multi(const DualDev &D, const ClusterDev &C, int nr, const double *yb, int64_t ystr, double *t,
                            int64_t tstr, hipStream_t st);
void launch_cluster_alpha(const ClusterDev &C, const double *e0s, double *lam0, double *alphaK, hipStream_t st);
void launch_cluster_apply_multi(const DualDev &D, const ClusterDev &C, int nr, const double *lam0, int64_t lstr,
                                const double *alphaK, int64_t astr, int addR, double *yb, int64_t ystr,
                                hipStream_t st);
// set-up: F0 (into C.F0inv) from the corner blocks; H = F0^-1 G0^T and the
// regularised Sa (into Sa) once C.F0inv holds the inverse
void launch_cluster_setup_f0(const ClusterDev &C, hipStream_t st);
void launch_cluster_setup_sa(const ClusterDev &C, double *Sa, hipStream_t st);
// alpha_tot(s) = alphaK(s) (hybrid) + N_s alpha_coarse
void launch_alpha_combine(const DualDev &D, const ClusterDev &C, bool hybrid, const double *alpha_coarse,
                          double *alpha_tot, hipStream_t st);
void launch_dot(const double *a, const double *b, int64_t n, double *partial, double *out, hipStream_t st);
void launch_p_update(double *p, const double *y, int64_t n, const double *num, const double *den, int first,
                     hipStream_t st);
void launch_xr_update(double *x, double *r, const double *p, const double *Ap, int64_t n, const double *num,
                      const double *den, hipStream_t st);
void launch_axpby(double *y, const double *a, const double *b, int64_t n, double alpha, double beta, hipStream_t st);
void launch_copy_scalar(const double *src, double *dst, hipStream_t st);
void launch_symmetrize(double *A, int n, hipStream_t st);
void launch_primal(const PrimalDev &P, int nsub, const double *ub, double *u, hipStream_t st);
// new loads f (local DOF order, all subdomains): P.yf = L^-1 f_int, hf = f_b - K_bi K_ii^-1 f_int
void launch_rhs_prepare(const PrimalDev &P, int nsub, const double *f, double *hf, hipStream_t st);
// nr load vectors at once (strides between vectors fstr, hstr, ystr; work:
// 16 values per interior row, P.work_off layout)
void launch_rhs_prepare_multi(const PrimalDev &P, int nsub, int nr, const double *f, int64_t fstr, double *hf,
                              int64_t hstr, double *yf, int64_t ystr, double *work, hipStream_t st);
void launch_primal_multi(const PrimalDev &P, int nsub, int nr, const double *ub, int64_t bstr, const double *yf,
                         int64_t ystr, const double *alpha, int64_t astr, double *u, int64_t ustr, double *work,
                         hipStream_t st);
// sens[e] = -w_e^T Ke0 u_e over the NE[0] x NE[1] x NE[2] HEXA8 grid (nodal fields, ndof nd)
void launch_element_sensitivity(const int *NE, int nd, const double *Ke0, const double *u, const double *w,
                                double *sens, hipStream_t st);
// structured grid: K_i CSR values on the device (box: 6 ints per subdomain,
// first global element and elements per direction; NE: global elements)
void launch_grid_stiffness(int nsub, int nd, const int *box, const int *NE, const double *Ke, const double *escale,
                           const int *rowptr, const int64_t *rp_off, const int64_t *nz_off, int max_nnz, double *vals,
                           hipStream_t st);
// dst[i] = src[pos[i]]
void launch_gather_values(int64_t n, const int64_t *pos, const double *src, double *dst, hipStream_t st);
// plain copy of n doubles (n multiple of 4): calibration of HBM byte counters
void launch_copy(const double *src, double *dst, int64_t n, hipStream_t st);
// read-only stream over n doubles (n even) with `wgs` workgroups (out: wgs doubles)
void launch_stream_read(const double *src, int64_t n, int wgs, double *out, hipStream_t st);

// SUPER_DIRICHLET: yb = K_bb xb - K_bn diag(K_nn)^-1 K_nb xb (n: interior DOFs
// next to the boundary; t holds n_near values of scratch)
void launch_superdir(const DualDev &D, int64_t n_near, const int64_t *sib_ptr, const int64_t *sib_col,
                     const double *sib_val, const double *sdiag, const int64_t *sbi_ptr, const int64_t *sbi_col,
                     const double *sbi_val, double *t, const double *xb, double *yb, hipStream_t st);
// KKT sums per subdomain (out: 2 per subdomain) and the boundary values of u
// (ub); part holds 2 nsub kkt_groups(max_n) partial sums (max_n: the largest
// subdomain's DOF count)
int kkt_groups(int max_n);
void launch_kkt_primal(const PrimalDev &P, int nsub, int max_n, const double *u, const double *f, const double *bt1,
                       const double *bt0, double *ub, double *part, double *out, hipStream_t st);
// bt0 += B0^T lambda0 (C.lambda0) on the boundary (bt0 cleared by the caller)
void launch_b0t(const DualDev &D, const ClusterDev &C, double *bt0, hipStream_t st);
// u <- trunc(u D) / D (FETISolver::solve output truncation); partial: 256 doubles
void launch_truncate(double *u, int64_t n, double precision, double *partial, hipStream_t st);

}  // namespace esp
