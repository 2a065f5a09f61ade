/*
 * espreso_amd -- MI355X-native Total FETI dual-loop core, C ABI.
 *
 * The reference's C++ solver layer (src/solver, src/linearsolver) reaches the
 * accelerator through these entry points.  Each replaces one step of the
 * reference's FETISolver / IterSolver / Cluster call chain:
 *
 *   espreso_feti_create          FETISolver::update -> new SuperCluster/IterSolver
 *                                (src/solver/generic/FETISolver.cpp:59-76)
 *   espreso_feti_set_subdomain   Domain::SetDomain inputs: K_i, f_i, R_i (kernel),
 *                                B1 rows, B1_scale_vec (src/solver/generic/Domain.cpp:67-215,
 *                                src/solver/generic/FETISolver.cpp:99-103)
 *   espreso_feti_set_dual        B1c (vec_c) (src/solver/specific/cluster.cpp:2732-2769)
 *   espreso_feti_setup           FETISolver::init: G/GG^T (itersolver.cpp:4931-5121),
 *                                Dirichlet preconditioner (clustercpu.cpp:167-333),
 *                                local Schur complements (clustercpu.cpp:11-40)
 *   espreso_feti_solve           FETISolver::solve -> Solve_RegCG + primal
 *                                (FETISolver.cpp:119-149, itersolver.cpp:1406-1724, :372-431)
 *   espreso_feti_apply_F         IterSolverCPU::apply_A_l_comp_dom_B (itersolvercpu.cpp:9-114)
 *   espreso_feti_apply_prec      IterSolverCPU::Apply_Prec (itersolvercpu.cpp:520-602)
 *   espreso_feti_project         IterSolverBase::Projector_Inv mode 0 (itersolver.cpp:5541-5623)
 *
 * plus a structured-grid problem builder standing in for the reference's
 * GRID generator + assembler (gridgenerator.cpp:44-69, physics.cpp:147-228,
 * equalityconstraints.cpp:25-423, structuralmechanics3d.cpp:34-184).
 *
 * Conventions: every function returns 0 on success or a negative status; the
 * message of the last failure on the calling thread is espreso_last_error().
 * espreso_feti_solve / _resolve return ESPRESO_NOT_CONVERGED (positive) when
 * the stop test was not met within max_iterations (or the mixed-precision
 * refinement limit); the stats and the last iterate are filled in as usual.
 * All pointers are caller-owned host memory unless the name says "device".
 * The library never calls exit().
 */
#ifndef ESPRESO_AMD_H
#define ESPRESO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ESPRESO_OK = 0,
    ESPRESO_ERR_ARG = -1,        /* invalid argument                     */
    ESPRESO_ERR_HIP = -2,        /* HIP runtime failure / no device      */
    ESPRESO_ERR_NUMERIC = -3,    /* factorisation breakdown, NaN/Inf     */
    ESPRESO_ERR_STATE = -4,      /* call out of order                    */
    ESPRESO_ERR_MEMORY = -5,     /* device memory budget exceeded        */
    ESPRESO_NOT_CONVERGED = 1    /* solved, stop test not met (stats filled) */
};

/* Structured box problem (same field meaning as the oracle's). */
typedef struct {
    int32_t physics;            /* 0: 3D linear elasticity, 1: 3D heat        */
    int32_t clusters[3];
    int32_t domains[3];         /* subdomains per cluster                     */
    int32_t elements[3];        /* elements per subdomain                     */
    double  length[3];
    double  young, poisson, density;
    double  accel[3];
    double  conductivity;
    double  heat_source;
    int32_t fixed[6];           /* faces x0,x1,y0,y1,z0,z1: fixed component bitmask */
    double  fixed_value[6];
    double  pressure[6];        /* normal pressure (>0 pushes inward)          */
    uint64_t traction_seed;
    double  traction_scale;
    uint64_t young_seed;
    uint64_t conductivity_seed;
    const double *material_elem; /* optional, per element (index ex + NEx*(ey + NEy*ez)):
                                    Young's modulus (elasticity) or conductivity (heat);
                                    overrides young/young_seed/conductivity(_seed)       */
    const int32_t *slab_elements; /* optional uneven subdomains: elements of every
                                    subdomain slab, x slabs then y then z (NSx+NSy+NSz
                                    values); overrides elements[]                      */
    int32_t device_assembly;    /* espreso_grid_load: 1 = the K_i values are computed on
                                   the GPU from the unit element matrix and the element
                                   materials (the host builds structure, loads, kernels
                                   and B1/B0 only); 0 = host-assembled values         */
    int32_t b0_type;            /* HYBRID_FETI B0: 0 = CORNERS (corner nodes, one row per
                                   DOF), 1 = KERNELS (per face pair of a cluster, one row
                                   per kernel column of the master subdomain; the
                                   reference's default, src/config/ecf/solver/feti.cpp:90) */
} espreso_grid_problem;

/* Values follow the reference enums (src/config/ecf/solver/feti.h). */
enum {
    ESPRESO_PREC_NONE = 0,
    ESPRESO_PREC_LUMPED = 1,          /* K_bb                                       */
    ESPRESO_PREC_WEIGHT = 2,          /* B B^T                                      */
    ESPRESO_PREC_DIRICHLET = 3,       /* K_bb - K_bi K_ii^-1 K_ib (dense S_D)      */
    ESPRESO_PREC_SUPER_DIRICHLET = 4  /* K_bb - K_bi diag(K_ii)^-1 K_ib (sparse,
                                         clustercpu.cpp:243-314, itersolvercpu.cpp:547-550) */
};
enum { ESPRESO_SC_FP64 = 0, ESPRESO_SC_FP32 = 1, ESPRESO_SC_BF16 = 2 };

typedef struct {
    int32_t method;             /* 0 TOTAL_FETI, 1 HYBRID_FETI                 */
    int32_t preconditioner;     /* ESPRESO_PREC_*                              */
    int32_t max_iterations;
    int32_t scaling;            /* 1: K-diagonal B1 scaling; 0: 1/multiplicity */
    int32_t sc_precision;       /* storage of F_i: ESPRESO_SC_*                */
    int32_t device;             /* HIP device ordinal                          */
    double  precision;          /* stop: ||P r|| < precision*min(||P r0||,||b||) */
    double  memory_budget_gb;   /* 0 = 85% of free device memory              */
    int32_t sd_precision;       /* storage of the Dirichlet S_D: ESPRESO_SC_*, any
                                   of the three whatever sc_precision is (S_D is the
                                   preconditioner: it changes the iteration count,
                                   not the converged solution)                    */
    int32_t truncate_output;    /* 1: truncate the primal solution to the digits the
                                   stop test guarantees, u <- trunc(u D)/D with
                                   D = 1/precision/max(1, 10^ceil(log10 max|u|)), as
                                   FETISolver::solve does (FETISolver.cpp:131-148);
                                   0 (default): the solution as computed          */
} espreso_feti_options;

typedef struct {
    int32_t iterations;
    int32_t n_lambda;
    int32_t n_coarse;
    int32_t n_subdomains;
    double  final_norm;
    double  tol;
    double  norm_b;
    double  time_assembly;      /* S_i, S_D assembly on the GPU (s)            */
    double  time_coarse;        /* G, GG^T setup (s)                           */
    double  time_pcg;           /* PCG loop (s)                                */
    double  time_primal;        /* primal recovery (s)                         */
    double  time_total;         /* setup + solve (s)                           */
    double  sc_bytes;           /* bytes of packed F_i (storage precision)     */
    double  sd_bytes;           /* bytes of packed S_D                         */
    double  assembly_flops;     /* flops executed by the assembly kernels      */
    int32_t refinement_steps;   /* outer FP64 refinement steps (mixed precision) */
    int32_t pad;
    double  time_load;          /* host preprocessing + upload of the inputs (s), not in time_total */
    double  symv_bytes;         /* algorithmic bytes of one F SYMV launch: packed F_i in storage
                                   precision + lambda gather (8+4+8 B per B1 entry) + y_b (8 B) */
    double  symv_bytes_sd;      /* the same for one S_D SYMV launch (FP64)                    */
    int32_t converged;          /* 1: final_norm < tol; 0: stopped at the iteration limit  */
    int32_t pad2;
    /* KKT report of the primal/dual solution (itersolver.cpp:171-370), computed on the
     * device after the primal recovery:
     *   kkt_primal = ||K u - f + B1^T lambda (+ B0^T lambda0)|| / ||f - B1^T lambda||
     *                over all subdomain DOFs (K the unregularised K_i),
     *   kkt_dual   = ||B1 u - c|| / norm_c with norm_c = ||c||, plus 1 when ||c|| < 1
     *                (the reference's normalisation of the equality constraints).   */
    double  kkt_primal;
    double  kkt_dual;
    double  time_build;         /* espreso_grid_load only: the structured-grid builder on the
                                   host (the reference's assembler), before set_subdomain  */
} espreso_feti_stats;

typedef struct espreso_feti_s espreso_feti;

const char *espreso_last_error(void);
int espreso_version(void);

/* ---- problem builder (host) ----------------------------------------------
 * Builds the decomposed problem of a structured box and feeds it to a FETI
 * instance: per-subdomain K_i (CSR), f_i, orthonormal kernel R_i, B1 rows
 * (Dirichlet + redundant gluing) with scaling, and c.  Lambda order: all
 * Dirichlet rows (by global DOF, then subdomain), then gluing rows (by global
 * DOF, then subdomain pair).                                                  */
int espreso_grid_sizes(const espreso_grid_problem *p, int32_t *n_subdomains,
                       int32_t *n_nodes, int32_t *ndof);
int espreso_grid_load(espreso_feti *h, const espreso_grid_problem *p);
/* Host-only export of subdomain i of the builder (no device needed).
 * sizes: [0] n, [1] nnz, [2] kdim, [3] lambda rows of i, [4] total lambda,
 * [5] subdomains.  Pass NULL arrays to query sizes only.  c (total lambda)
 * may be NULL.  With p->device_assembly the builder runs as it does for the
 * device assembly (structure, loads, kernels, B1 only) and values are 0.    */
int espreso_grid_export(const espreso_grid_problem *p, int32_t scaling, int32_t i, int32_t *sizes,
                        int32_t *rowptr, int32_t *colidx, double *values, double *f, double *R,
                        int32_t *lambda, int32_t *dof, double *sign, double *scale, double *c);
/* Host-only export of the hybrid-FETI data of subdomain i: its cluster and
 * its B0 corner entries (cluster-local row, local DOF, +1/-1).  *n0 returns
 * the entry count; pass NULL arrays to query it.                             */
int espreso_grid_export_cluster(const espreso_grid_problem *p, int32_t i, int32_t *cluster, int32_t *n0,
                                int32_t *row, int32_t *dof, double *sign);
/* Unit-modulus element matrix K_e0 of the grid (ndof*8 x ndof*8, row-major,
 * DOFs node-major in the reference's corner order): K_e = E_e K_e0.          */
int espreso_grid_unit_element(const espreso_grid_problem *p, double *Ke);
/* global node id of every local node of subdomain i (after espreso_grid_load) */
int espreso_grid_l2g(espreso_feti *h, int32_t i, int32_t *l2g);
/* merge per-subdomain primal of the last solve into nodal u (average of copies) */
int espreso_grid_nodal_solution(espreso_feti *h, double *u_nodal);
/* Assembled nodal load of the current subdomain loads (sum over copies). */
int espreso_grid_nodal_load(espreso_feti *h, double *g_nodal);
/* New nodal load g (n_nodes x ndof) for espreso_feti_resolve: each node's load
 * is split evenly over the subdomains that share it.                         */
int espreso_grid_set_nodal_rhs(espreso_feti *h, const double *g_nodal);
/* Block re-solve of nodal load cases on the grid instance: g holds nrhs
 * nodal loads (n_nodes x ndof each) split over the subdomain copies of a node
 * as espreso_grid_set_nodal_rhs does; u_nodal receives the nodal solutions
 * averaged over the copies as espreso_grid_nodal_solution does.  The solve is
 * espreso_feti_resolve_block (iterations, final_norms, stats may be NULL).  */
int espreso_grid_resolve_block(espreso_feti *h, int32_t nrhs, const double *g, double *u_nodal, int32_t *iterations,
                               double *final_norms, espreso_feti_stats *stats);
/* New per-element material (Young's modulus or conductivity, index
 * ex + NEx*(ey + NEy*ez)) on a loaded grid instance: the K_i values are
 * recomputed on the GPU from the unit element matrix and the next
 * espreso_feti_setup / espreso_feti_solve use them
 * (espreso_feti_update_values for every subdomain).                          */
int espreso_grid_update_material(espreso_feti *h, const double *material_elem);
/* Element sensitivities on the GPU: sens[e] = -w_e^T K_e0 u_e for nodal
 * fields u, w (n_nodes x ndof); with w = u (the compliance adjoint) this is
 * d(f^T u)/dE_e.  No counterpart in the reference (adjoint extension).       */
int espreso_grid_element_sensitivity(espreso_feti *h, const double *u_nodal, const double *w_nodal, double *sens);

/* ---- FETI core --------------------------------------------------------- */
int espreso_feti_create(espreso_feti **h, int32_t n_subdomains, const espreso_feti_options *o);
int espreso_feti_destroy(espreso_feti *h);
/* sizes[0] = subdomains, [1] = lambda (after set_dual, else -1),
 * [2] = coarse dimension (after setup, else 0)                               */
int espreso_feti_sizes(espreso_feti *h, int32_t *sizes);

/* K: full symmetric CSR (0-based) of n x n; f: n; R: n x kdim column-major,
 * orthonormal columns spanning ker K; lambda rows touching the subdomain
 * (ascending lambda id) with local DOF, sign (+1/-1) and scaling weight.
 * K is read as symmetric: when every row's columns are sorted (n <= 65536),
 * only its lower triangle (column <= row) is kept and the upper entries take
 * their mirrors' values; espreso_feti_setup then returns ESPRESO_ERR_ARG if
 * an upper entry's mirror is missing from the structure.  The arrays are
 * copied (into pinned host memory where possible); the caller may free them. */
int espreso_feti_set_subdomain(espreso_feti *h, int32_t i, int32_t n,
                               const int32_t *rowptr, const int32_t *colidx, const double *values,
                               const double *f, int32_t kdim, const double *R,
                               int32_t nlam, const int32_t *lambda, const int32_t *dof,
                               const double *sign, const double *scale);
/* c (n_lambda values).  Before the first espreso_feti_setup any size; after
 * it only the loaded size (new Dirichlet values: b is recomputed and the next
 * espreso_feti_resolve uses them).                                           */
int espreso_feti_set_dual(espreso_feti *h, int32_t n_lambda, const double *c);
/* New values of K_i on the structure of espreso_feti_set_subdomain (same
 * rowptr/colidx, nnz values), for nonlinear and time-stepping callers.  The
 * next espreso_feti_setup re-assembles F_i and S_D from them and keeps the
 * boundary sets, lambda maps, coarse space G, (G G^T)^-1 and the workspace;
 * the kernel R_i and the B1 scaling given at set_subdomain stay in force, so
 * R_i must still span ker K_i (it does when only material values change).
 * The reference rebuilds its whole SuperCluster on a K update
 * (FETISolver::update with Matrices::K, FETISolver.cpp:59-76).               */
int espreso_feti_update_values(espreso_feti *h, int32_t i, const double *values);
/* Replace the load vector of subdomain i (local DOF order) after setup; the
 * next espreso_feti_resolve solves with it, reusing F_i, S_D, the interior
 * factors and the coarse space.  FETISolver::update with Matrices::f only
 * (FETISolver.cpp:59-116).                                                    */
int espreso_feti_set_rhs(espreso_feti *h, int32_t i, const double *f);
/* Solve again with the loads set by espreso_feti_set_rhs (no reassembly):
 * new b and e, projected PCG, primal recovery.                               */
int espreso_feti_resolve(espreso_feti *h, espreso_feti_stats *stats);
/* Block re-solve: nrhs load cases on the assembled operators, solved
 * together -- one projected PCG per case (own scalars and stop test) in
 * lockstep, so that each F and Dirichlet apply reads the F_i / S_D bytes once
 * for a group of up to 16 cases.  f: nrhs consecutive load vectors, each the
 * subdomain-local loads of all subdomains in subdomain order (n_i values of
 * subdomain i, as espreso_feti_set_rhs takes them); u: the primal solutions
 * in the same layout; iterations and final_norms (per case) and stats may be
 * NULL (stats: maximum iterations and norm over the cases, times of the load
 * preparation in time_coarse, of the PCG and of the primal recovery).  The
 * instance's own loads are left as they were.  With low-precision F_i or S_D
 * storage the cases are solved one after another.  No counterpart in the
 * reference, which solves one Matrices::f per FETISolver::solve
 * (FETISolver.cpp:59-149, itersolver.cpp:1406-1724).                        */
int espreso_feti_resolve_block(espreso_feti *h, int32_t nrhs, const double *f, double *u, int32_t *iterations,
                               double *final_norms, espreso_feti_stats *stats);
/* sum over subdomains of f_i^T u_i for the last solution (the compliance
 * f^T u of the assembled problem)                                            */
int espreso_feti_load_work(espreso_feti *h, double *c);
/* Hybrid FETI (options.method = 1): the cluster of subdomain i and its B0
 * entries (cluster-local row id, local DOF, value), any number per row; the
 * subdomains of a cluster are numbered consecutively, at most 256 B0 rows
 * touch one subdomain, and every row has entries.  Corner rows (one +1 and
 * one -1 entry: Constraints::B0Corners, equalityconstraints.cpp:504-552) and
 * kernel rows (R_master-weighted face DOFs of two subdomains:
 * Constraints::B0Kernels, equalityconstraints.cpp:425-502) both fit.
 * Replaces the ClusterBase::CompressB0 (cluster.cpp:1186-1224) inputs.       */
int espreso_feti_set_cluster(espreso_feti *h, int32_t i, int32_t cluster, int32_t n0, const int32_t *row,
                             const int32_t *dof, const double *sign);

/* GPU setup: boundary Schur complements, F_i = (K^+)_bb, S_D, coarse problem.
 * The first call also preprocesses and uploads the inputs; later calls redo
 * only the GPU part on the resident inputs.                                  */
int espreso_feti_setup(espreso_feti *h);

/* Projected PCG to the stop criterion, then primal recovery. */
int espreso_feti_solve(espreso_feti *h, espreso_feti_stats *st);
/* per-subdomain primal solution of the last solve (n_i values) */
int espreso_feti_primal(espreso_feti *h, int32_t i, double *u);
/* dual solution lambda of the last solve (n_lambda values) */
int espreso_feti_dual(espreso_feti *h, double *lambda);

/* Dual operators on host vectors of length n_lambda (parity / inspection). */
int espreso_feti_apply_F(espreso_feti *h, const double *x, double *y);
int espreso_feti_apply_prec(espreso_feti *h, const double *x, double *y);
int espreso_feti_project(espreso_feti *h, const double *x, double *y);
/* y = P x (as espreso_feti_project) on device vectors x, y of length n_lambda,
 * launched on `stream` (a hipStream_t; NULL: the instance's stream) without
 * waiting for it.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * It allocates nothing: its scratch (G x per subdomain and per coarse block,
 * (G G^T)^-1 G x: sum of the kernel dimensions + 2 x the coarse dimension
 * doubles) is a static device array of 2^20 doubles (8 MiB) per GPU, shared by
 * every call on that GPU, so calls that overlap on one GPU race.  A larger
 * coarse space is refused with ESPRESO_ERR_ARG. */
int espreso_feti_project_device(espreso_feti *h, const double *x, double *y, void *stream);
int espreso_feti_vec_b(espreso_feti *h, double *b);
/* SYMV work-item sizes in matrix elements per workgroup, targets[3]: FP64,
 * FP32 and BF16 storage (0 where neither F_i nor S_D is stored in that
 * precision).  The first espreso_feti_setup picks them by timing candidates
 * on the assembled operators (cached per process by storage precision and
 * subdomain interface size); ESPRESO_SYMV_ITEM fixes them and
 * ESPRESO_AUTOTUNE=0 keeps the default (327680).                            */
int espreso_feti_symv_items(espreso_feti *h, int64_t *targets);
/* K_i CSR values of subdomain i as the device holds them (after setup; the
 * CSR of espreso_feti_set_subdomain, or of the grid builder with
 * device_assembly), for parity checks of the device assembly              */
int espreso_feti_subdomain_values(espreso_feti *h, int32_t i, double *vals);
/* dense m_i x m_i (column-major, boundary DOFs in ascending local order):
 * which 0: F_i = (K^+)_bb in FP64, 1: S_D, 2: F_i as stored (after rounding) */
int espreso_feti_subdomain_dense(espreso_feti *h, int32_t i, int32_t which, double *out);
int espreso_feti_subdomain_boundary(espreso_feti *h, int32_t i, int32_t *m, int32_t *bdofs);

/* Timed dual-operator applies on device-resident vectors (benchmark use):
 * runs `reps` applies of F (op 0) or of the preconditioner (op 1) and
 * returns the mean time of one apply and of its SYMV kernel alone in ms
 * (HIP events recorded on the stream the kernels run on).                   */
int espreso_feti_time_apply(espreso_feti *h, int32_t op, int32_t reps,
                            double *ms_apply, double *ms_symv);

/* Diagnostics of the FP64 MFMA tile routine used by the assembly.
 * espreso_debug_mfma: one v_mfma_f64_16x16x4f64 with lane l holding a[l], b[l];
 * out[4*l + v] = accumulator value v of lane l; then one v_mfma_f64_4x4x4f64
 * (4 blocks) on the same operands, out[256 + l] = its result in lane l.
 * espreso_debug_tile_gemm: C = alpha op(A) op(B) + beta C for one tile
 * (M, N <= 64), column-major host arrays.                                     */
int espreso_debug_mfma(const double *a, const double *b, double *out);
/* Plain device copy of `bytes` (read + write), `reps` times: mean ms per copy.
 * Used to calibrate HBM byte counters of the profiler.                      */
int espreso_debug_copy(int64_t bytes, int32_t reps, double *ms);
/* Read-only stream over `bytes` with `workgroups` workgroups of 256 threads,
 * `reps` times: mean ms per pass (the achievable HBM read rate that the
 * bandwidth-bound kernels are compared with).                               */
int espreso_debug_stream(int64_t bytes, int32_t reps, int32_t workgroups, double *ms);
/* FP64 MFMA issue rate with register operands only (4096 workgroups of 4 waves,
 * `chains` = 4, 8, 16 or 32 independent accumulators per wave on one operand
 * pair): shape 0 v_mfma_f64_16x16x4f64, 1 v_mfma_f64_4x4x4f64 (4 blocks);
 * shapes 2 (4x4x4) and 3 (16x16x4): `chains` A and B operand registers and
 * chains x chains accumulators, as in a register-blocked tile; TFLOP/s.  The
 * ceiling that the assembly kernels' MFMA rate is compared with.
 * workgroups_per_cu > 0 caps the workgroups sharing a CU (through unused LDS),
 * i.e. the waves per SIMD; 0 leaves the occupancy to the register use.      */
int espreso_debug_mfma_rate(int32_t shape, int32_t chains, int32_t workgroups_per_cu, double *tflops);
int espreso_debug_tile_gemm(int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const double *A, int32_t lda,
                            const double *B, int32_t ldb, double *C, int32_t ldc, double alpha, double beta);

#ifdef __cplusplus
}
#endif
#endif
