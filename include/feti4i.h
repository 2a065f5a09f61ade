/*
 * FETI4I -- the ESPRESO library interface, served by the MI355X FETI core.
 *
 * Drop-in for the reference's include/feti4i.h: the same types, functions,
 * argument meanings and option enums (reference include/feti4i.h:17-270).
 * Differences, all additive:
 *   - no "mpi.h": one process drives one cluster on one GPU, so
 *     FETI4IMPIInt is a plain int and `neighbours` must be empty;
 *   - errors never terminate the process: a failing call records a status
 *     and message, readable with FETI4IGetLastStatus / FETI4IGetLastError;
 *   - a few inspection calls (FETI4IInstanceInfo, FETI4IExportSubdomain,
 *     FETI4IGetStats) at the end of this header.
 * The shared library is built with FETI4I_INT_WIDTH 32.
 */
#ifndef FETI4I_H_
#define FETI4I_H_

#ifndef FETI4I_INT_WIDTH
#define FETI4I_INT_WIDTH 32
#endif
#ifndef FETI4I_REAL_WIDTH
#define FETI4I_REAL_WIDTH 64
#endif

#if FETI4I_INT_WIDTH == 32
typedef int FETI4IInt;
#elif FETI4I_INT_WIDTH == 64
typedef long FETI4IInt;
#else
#error "FETI4I_INT_WIDTH must be 32 or 64"
#endif

typedef int FETI4IMPIInt;

#if FETI4I_REAL_WIDTH == 64
typedef double FETI4IReal;
#else
#error "FETI4I_REAL_WIDTH must be 64"
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* opaque handles */
typedef struct FETI4IStructMatrix *FETI4IMatrix;
typedef struct FETI4IStructInstance *FETI4IInstance;

/* Default option arrays (sizes FETI4I_INTEGER_OPTIONS_SIZE / FETI4I_REAL_OPTIONS_SIZE). */
void FETI4ISetDefaultIntegerOptions(FETI4IInt *options);
void FETI4ISetDefaultRealOptions(FETI4IReal *options);

/* Empty stiffness matrix; element indices given later are shifted by -indexBase. */
void FETI4ICreateStiffnessMatrix(FETI4IMatrix *matrix, FETI4IInt type, FETI4IInt indexBase);

/* Append one symmetric element matrix: `values` is dofsSize x dofsSize,
 * row-major, over the local DOFs `dofs`; `nodes` lists the element nodes. */
void FETI4IAddElement(FETI4IMatrix matrix, FETI4IInt type, FETI4IInt nodesSize, FETI4IInt *nodes,
                      FETI4IInt dofsSize, FETI4IInt *dofs, FETI4IReal *values);

/* Problem instance: `size` local DOFs with right-hand side `rhs` and the
 * local-to-global DOF map `l2g`; Dirichlet conditions on local DOFs.  The
 * mesh is split into integer_options[FETI4I_SUBDOMAINS] subdomains. */
void FETI4ICreateInstance(FETI4IInstance *instance, FETI4IMatrix matrix, FETI4IInt size, FETI4IReal *rhs,
                          FETI4IInt *l2g, FETI4IMPIInt neighbours_size, FETI4IMPIInt *neighbours,
                          FETI4IInt dirichlet_size, FETI4IInt *dirichlet_indices, FETI4IReal *dirichlet_values,
                          FETI4IInt *integer_options, FETI4IReal *real_options);

/* Solve; writes `solution_size` (= size) values: the DOF values averaged
 * over the subdomain copies of each DOF. */
void FETI4ISolve(FETI4IInstance instance, FETI4IInt solution_size, FETI4IReal *solution);

void FETI4IUpdateStiffnessMatrix(FETI4IInstance instance, FETI4IMatrix stiffnessMatrix);
void FETI4IUpdateRhs(FETI4IInstance instance, FETI4IInt size, FETI4IReal *values);
void FETI4IUpdateDirichlet(FETI4IInstance instance, FETI4IInt size, FETI4IInt *indices, FETI4IReal *values);

/* Release a matrix or an instance. */
void FETI4IDestroy(void *ptr);

/* ---- additions of this library ------------------------------------------ */
/* 0 when the last FETI4I call on this thread succeeded, else a negative code */
int FETI4IGetLastStatus(void);
const char *FETI4IGetLastError(void);
/* info (8 entries): [0] subdomains, [1] lambda rows, [2] coarse dimension,
 * [3] last PCG iterations, [4] DOFs per node detected, [5] max kernel
 * dimension, [6] B0 corner rows (HYBRID_FETI, else 0), [7] FETI method     */
int FETI4IInstanceInfo(FETI4IInstance instance, FETI4IInt *info);
/* Decomposed data of subdomain i as handed to the FETI core; sizes[0] n,
 * [1] nnz, [2] kernel dimension, [3] lambda rows of i.  NULL arrays query
 * sizes only.  gdofs: instance-local DOF of every subdomain DOF.            */
int FETI4IExportSubdomain(FETI4IInstance instance, FETI4IInt i, FETI4IInt *sizes, FETI4IInt *gdofs,
                          FETI4IInt *rowptr, FETI4IInt *colidx, FETI4IReal *values, FETI4IReal *f,
                          FETI4IReal *R, FETI4IInt *lambda, FETI4IInt *dof, FETI4IReal *sign,
                          FETI4IReal *scale);
/* HYBRID_FETI corners of subdomain i (the instance is one cluster):
 * *size entries of (cluster B0 row, subdomain-local DOF, +1/-1); NULL arrays
 * query the size only.  Every row has one +1 and one -1 entry, in two
 * subdomains, on copies of the same instance DOF.                           */
int FETI4IExportCorners(FETI4IInstance instance, FETI4IInt i, FETI4IInt *size, FETI4IInt *row, FETI4IInt *dof,
                        FETI4IReal *sign);
/* stats[0] iterations, [1] final ||P r||, [2] tolerance, [3] assembly s,
 * [4] PCG s, [5] total GPU s                                                */
int FETI4IGetStats(FETI4IInstance instance, FETI4IReal *stats);

#ifdef __cplusplus
}
#endif

/* integer options (reference feti4i.h:220-236) */
typedef enum {
    FETI4I_SUBDOMAINS,
    FETI4I_MAX_ITERATIONS,
    FETI4I_FETI_METHOD,
    FETI4I_PRECONDITIONER,
    FETI4I_CGSOLVER,
    FETI4I_N_MICS,
    FETI4I_VERBOSE_LEVEL,
    FETI4I_MEASURE_LEVEL,
    FETI4I_PRINT_MATRICES,
    FETI4I_SC_SIZE,
    FETI4I_INTEGER_OPTIONS_SIZE
} FETI4IIntegerOptions;

/* real options (reference feti4i.h:244-248) */
typedef enum {
    FETI4I_PRECISION,
    FETI4I_REAL_OPTIONS_SIZE
} FETI4IRealOptions;

/* matrix types (reference feti4i.h:254-258) */
typedef enum {
    FETI4I_REAL_SYMMETRIC_POSITIVE_DEFINITE,
    FETI4I_REAL_SYMMETRIC_INDEFINITE,
    FETI4I_REAL_UNSYMMETRIC
} FETI4IMatrixType;

/* element types (reference feti4i.h:265-270) */
typedef enum {
    FETI4I_POINT,
    FETI4I_LINE,
    FETI4I_PLANE,
    FETI4I_VOLUME
} FETI4IElementType;

#endif /* FETI4I_H_ */
