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
    int64_t *foff_lp = nullptr;      // offsets of the low-precision copy
    double *SD = nullptr;
    double *GGinv = nullptr;
    // SYMV work lists: FP64 geometry (F_i and S_D) and low-precision geometry
    SymvItem *items_d = nullptr;  int n_items_d = 0;  int *item0_d = nullptr;  int *nitems_d = nullptr;
    SymvItem *items_lp = nullptr; int n_items_lp = 0; int *item0_lp = nullptr; int *nitems_lp = nullptr;
    double *ypart = nullptr;
    double *yb = nullptr, *xb = nullptr;  // nbnd scratch
    double *cd = nullptr, *ce = nullptr;  // coarse scratch
};

struct PrimalDev {
    int *ni = nullptr, *m = nullptr, *n = nullptr, *bw = nullptr, *kdim = nullptr, *coff = nullptr;
    int64_t *band_off = nullptr, *yf_off = nullptr, *perm_off = nullptr, *rp_off = nullptr, *nz_off = nullptr;
    int64_t *boff = nullptr, *work_off = nullptr, *r_off = nullptr;
    double *band = nullptr, *yf = nullptr, *vals = nullptr, *R = nullptr, *alpha = nullptr, *work = nullptr;
    int *perm = nullptr, *iperm = nullptr, *rowptr = nullptr, *colidx = nullptr;
};

// y_b (D.yb) = A x_b for every subdomain; A = S_D (sd) or F_i in storage
// precision `prec` (0 FP64, 1 FP32, 2 BF16).  x is lambda-space and gathered
// with sign (or sign*scale when `scaled`) if `gather`, else boundary-space.
// ev0/ev1 (optional) are recorded around the SYMV kernel alone.
void launch_symv(const DualDev &D, bool sd, int prec, const double *x, int gather, bool scaled, hipStream_t st,
                 hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void launch_lambda_gather(const DualDev &D, const double *yb, bool scaled, double *y, hipStream_t st);
void launch_boundary_gather(const DualDev &D, const double *x, bool scaled, double *xb, hipStream_t st);
void launch_kbb(const DualDev &D, const double *xb, double *yb, hipStream_t st);
// mode 0: y = x - G^T GGinv G x; mode 1: y = G^T GGinv x (x coarse); mode 2: y = GGinv G x (y coarse)
void launch_project(const DualDev &D, const double *x, double *y, int mode, hipStream_t st);
void launch_dot(const double *a, const double *b, int64_t n, double *partial, double *out, hipStream_t st);
void launch_p_update(double *p, const double *y, int64_t n, const double *num, const double *den, int first,
                     hipStream_t st);
void launch_xr_update(double *x, double *r, const double *p, const double *Ap, int64_t n, const double *num,
                      const double *den, hipStream_t st);
void launch_axpby(double *y, const double *a, const double *b, int64_t n, double alpha, double beta, hipStream_t st);
void launch_copy_scalar(const double *src, double *dst, hipStream_t st);
void launch_symmetrize(double *A, int n, hipStream_t st);
void launch_primal(const PrimalDev &P, int nsub, const double *ub, double *u, hipStream_t st);
// plain copy of n doubles (n multiple of 4): calibration of HBM byte counters
void launch_copy(const double *src, double *dst, int64_t n, hipStream_t st);

}  // namespace esp
