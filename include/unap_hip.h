/*
 * unap_hip.h -- C-ABI of the MI355X (gfx950) LDU solve path.
 *
 * Plain C types only (no HIP or torch types): opaque handles, int64 labels
 * (the reference's `label` under -DLABEL_INT64, SConstruct:52) and double
 * scalars.  Every call returns an int status (UNAP_HIP_OK = 0, negative on
 * error, message in unap_hip_last_error()) and never exits the process,
 * unlike the reference's CHECK_POINTER/ERROR_EXIT (src/base/unap.hpp:96-101).
 *
 * Vectors are in the caller's cell numbering.  `where` = UNAP_HIP_HOST for
 * host pointers, UNAP_HIP_DEVICE for device pointers (HBM-resident data, no
 * PCIe copy).  create() calls deep-copy their inputs, as the reference's
 * non-reUse constructors do (lduMatrix.cpp:28-68).
 *
 * Threading: one process-wide HIP stream; calls are not re-entrant (the
 * reference is not either).  One GPU (the current HIP device).
 */
#ifndef UNAP_HIP_H
#define UNAP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UNAP_HIP_OK 0
#define UNAP_HIP_ERR_ARG (-1)        /* bad argument */
#define UNAP_HIP_ERR_HIP (-2)        /* HIP runtime error (no device, OOM, ...) */
#define UNAP_HIP_ERR_FACE_ORDER (-3) /* sweep needs owner-sorted faces, owner < neighbour */
#define UNAP_HIP_ERR_STATE (-4)      /* handle used in a state that does not allow it */
#define UNAP_HIP_ERR_INTERNAL (-5)

#define UNAP_HIP_HOST 0
#define UNAP_HIP_DEVICE 1

typedef struct unap_hip_ldu_s* unap_hip_ldu;
typedef struct unap_hip_precond_s* unap_hip_precond;
typedef struct unap_hip_smoother_s* unap_hip_smoother;
typedef struct unap_hip_aggl_s* unap_hip_aggl;
typedef struct unap_hip_solver_s* unap_hip_solver;

/* Last error message of this thread ("" if none). */
const char* unap_hip_last_error(void);

/* Order of every FP64 dot product and norm (no reference counterpart; a parity
   aid).  UNAP_HIP_REDUCE_TREE (default): wave64 shuffle trees and fixed-order
   block partials, deterministic run to run.  UNAP_HIP_REDUCE_REFERENCE: one
   thread sums in the caller's cell order, exactly the serial loops of
   unapVectorOperations.cpp:21-30 and unapVector.hpp:478-489, so FP64 solves
   reproduce the reference bit for bit (slow: O(n) sequential per reduction). */
#define UNAP_HIP_REDUCE_TREE 0
#define UNAP_HIP_REDUCE_REFERENCE 1
int unap_hip_set_reduction_order(int order);

/* Solver iterations from the second on replayed from a HIP graph of one
   iteration (1, default) or launched one kernel at a time (0).  No reference
   counterpart; the results are identical either way.  The environment
   variable UNAP_HIP_GRAPHS=0 also turns the replay off. */
int unap_hip_set_graph_replay(int on);

/* Device name and compute-unit count of the current device. */
int unap_hip_device_info(char* name, int64_t name_len, int64_t* compute_units);

/* ------------------------------------------------------------- LDU matrix
 * Replaces lduMatrix(nCells, lowerAddr, upperAddr, lower, diag, upper)
 * (src/matrix/lduMatrix/lduMatrix.hpp:80-88, lduMatrix.cpp:28-68).  As there,
 * lower == upper (the same pointer) or lower == NULL makes the matrix
 * symmetric (one coefficient array).  Face f couples lowerAddr[f] (owner,
 * coefficient upper[f] in its row) and upperAddr[f] (neighbour, lower[f]).
 * Amul accepts any face order; the sweeps (DIC, DILU, Gauss-Seidel) need faces
 * sorted by owner with owner < neighbour, as the reference's do. */
int unap_hip_ldu_create(int64_t n_cells, int64_t n_faces, const int64_t* lower_addr,
                        const int64_t* upper_addr, const double* lower, const double* diag,
                        const double* upper, unap_hip_ldu* out);
int unap_hip_ldu_destroy(unap_hip_ldu A);
/* size() / nFaces / symm() (lduMatrix.hpp:210-231) and the level-schedule depth
   of the forward and backward sweeps (0 when faces are not owner-sorted).  The
   depths are computed on the first call that asks for them (a pass over every
   row); pass NULL for them to skip it. */
int unap_hip_ldu_info(unap_hip_ldu A, int64_t* n_cells, int64_t* n_faces, int* symmetric,
                      int64_t* fwd_levels, int64_t* bwd_levels);
/* lowerAddr() / upperAddr() / lower() / diag() / upper() (lduMatrix.hpp:122-129):
   host copies of the matrix's arrays (any pointer may be NULL; lower is upper
   for a symmetric matrix). */
int unap_hip_ldu_get(unap_hip_ldu A, int64_t* lower_addr, int64_t* upper_addr, double* lower, double* diag,
                     double* upper);
/* setMatrixCoeffients(diag, upper, lower) (lduMatrix.cpp:508-529): new values,
   same addressing; lower == upper or NULL keeps one coefficient array. */
int unap_hip_ldu_set_coeffs(unap_hip_ldu A, const double* diag, const double* upper,
                            const double* lower);
/* setMatrixTopology(upperAddr, lowerAddr) (lduMatrix.cpp:566-597): new
   addressing, same cell count.  As there, the coefficients are not touched:
   with the same face count they stay (now indexed by the new faces); with
   another face count the matrix refuses every use (UNAP_HIP_ERR_STATE) until
   unap_hip_ldu_set_coeffs gives values for the new faces.  Device layouts and
   cached agglomerations are rebuilt on next use; preconditioners built on the
   earlier addressing are refused (UNAP_HIP_ERR_STATE) where the reference's
   would run on stale data. */
int unap_hip_ldu_set_topology(unap_hip_ldu A, int64_t n_faces, const int64_t* lower_addr,
                              const int64_t* upper_addr);
/* spMV(Apsi, psi) (src/matrix/lduMatrix/lduMatrixSpMV.cpp:16-98). */
int unap_hip_ldu_spmv(unap_hip_ldu A, const double* psi, double* Apsi, int where);
/* No counterpart in the reference: unap_hip_ldu_spmv for device vectors that
   allocates nothing (no hipMalloc or hipFree, no allocating library call),
   for callers that cannot allocate (code under HIP graph capture, tight
   serving loops, harnesses that forbid it).  The internal-order staging lives
   in a fixed device scratch array, so the matrix may hold at most 2097152
   cells (UNAP_HIP_ERR_ARG beyond).  Needs the library context and the
   matrix's device layout, both built by one earlier unap_hip_ldu_spmv call on
   this matrix (UNAP_HIP_ERR_STATE otherwise).  Every call on one GPU shares
   that scratch: calls that overlap on one GPU race.  Synchronous, like
   unap_hip_ldu_spmv. */
int unap_hip_ldu_spmv_noalloc(unap_hip_ldu A, const double* psi, double* Apsi);

/* --------------------------------------------------------- preconditioners
 * Replaces matrix::preconditioner::precondition(w, r) (unapMatrix.hpp:337)
 * and its ldu implementations; type codes 1/2/3 are the reference's Fortran
 * precond codes (fortranInterface.cpp:227-238). */
#define UNAP_HIP_PRECOND_NONE 0 /* identity, unapMatrixPrecond.cpp:4-17 */
#define UNAP_HIP_PRECOND_DIAG 1 /* lduDiagPrecond.cpp:13-73 */
#define UNAP_HIP_PRECOND_DIC 2  /* lduDICPrecond.cpp:4-77 */
#define UNAP_HIP_PRECOND_DILU 3 /* lduDILUPrecond.cpp:4-86 */
#define UNAP_HIP_PRECOND_AMG 4  /* MGPreconditioner.hpp:11-49 (a stub there; see below) */

#define UNAP_HIP_SMOOTHER_GS 1    /* lduGaussSeidelSmoother.cpp:3-80 */
#define UNAP_HIP_SMOOTHER_CHEBY 2 /* chebySmoother.cpp:13-149 */

/* Agglomeration-AMG preconditioner: precondition(w, r) = { w = 0;
   MGSolver::solve(w, A, r) } running exactly n_vcycles V-cycles
   (MG.hpp:8-121, Vcycle.cpp:27-332) over the pairwise agglomeration of the
   face weights (|upper| when weights == NULL, MGPreconditioner.cpp:14-19).
   precision 64: the reference's arithmetic in FP64 (coarsest level solved by
   PCG+DIC / PBiCGStab+DILU to relTol 1e-2, solveCoarsestLevel.cpp:10-55).
   precision 32: FP32 hierarchy, smoothers and transfers, Chebyshev estimates
   taken once at setup, coarsest level solved directly in FP64 (a dense
   inverse by Gauss-Jordan elimination at setup); use it under the flexible
   solver UNAP_HIP_SOLVER_FCG. */
typedef struct {
    int64_t smoother;               /* UNAP_HIP_SMOOTHER_* (default GS) */
    int64_t n_cells_coarsest;       /* SET_nCellsInCoarsestLevel (default 50) */
    int64_t max_levels;             /* SET_maxLevels (default 50) */
    int64_t n_vcycles;              /* V-cycles per application (default 2, MGPreconditioner.cpp:11) */
    int64_t n_pre_sweeps;           /* MG.cpp:11-13 defaults 0 / 2 / 2 */
    int64_t n_post_sweeps;
    int64_t n_finest_sweeps;
    int64_t precision;              /* 64 or 32 */
    int64_t level_sweep_growth;     /* 1: coarse level l smooths nPost + l times (Vcycle.cpp:250,
                                       default); 0: every coarse level smooths nPost times */
    int64_t pair_steps;             /* pairwise agglomeration steps per hierarchy level, 1 (the
                                       reference, default) to 8: the hierarchy keeps levels
                                       pair_steps, 2 pair_steps, ... of the reference's pairwise
                                       agglomeration (counted from the finest) and its coarsest
                                       level, restriction maps composed */
    int64_t cycle;                  /* UNAP_HIP_CYCLE_*: the reference V-cycle (Chebyshev
                                       estimates on each solve's first call, coarsest level by
                                       PCG/PBiCGStab to 1e-2) or a fixed one (estimates once at
                                       setup, coarsest level by a dense FP64 inverse); default
                                       UNAP_HIP_CYCLE_AUTO: reference in FP64, fixed in FP32 */
    int64_t scaling;                /* correction scaling (scalingFactor.cpp, symmetric matrices):
                                       2 every level but the two coarsest (the reference,
                                       Vcycle.cpp:72,193; default), 1 the coarse levels only (not
                                       the finest), 0 none */
    double cheby_eig_ratio;         /* Chebyshev smoothing interval of every level:
                                       [lambda / cheby_eig_ratio, lambda * cheby_boost], lambda the
                                       level's eigenDiagPCG estimate.  Defaults 1.5 and 1.1, the
                                       reference's eigRatioCheby_ and boostFactorCheby_
                                       (chebySmoother.hpp:46-47, chebySmoother.cpp:63-64);
                                       cheby_eig_ratio > 1, cheby_boost > 0 */
    double cheby_boost;
    double cheby_eig_ratio_finest;  /* the finest level's ratio in place of cheby_eig_ratio;
                                       0 (default): cheby_eig_ratio on every level; else > 1.
                                       No counterpart in the reference (one ratio for all levels) */
} unap_hip_amg_params;
#define UNAP_HIP_CYCLE_AUTO (-1)
#define UNAP_HIP_CYCLE_REFERENCE 0
#define UNAP_HIP_CYCLE_FIXED 1

int unap_hip_amg_params_default(unap_hip_amg_params* p);
/* amg: the AMG parameters for UNAP_HIP_PRECOND_AMG (NULL: defaults).  For
   DIC and DILU only amg->precision is read (NULL: 64): 32 runs the sweeps in
   FP32 (FP32 coefficients and rD, the factorisation's rD computed in FP64 and
   rounded) as the preconditioner of an FP64 UNAP_HIP_SOLVER_FCG or
   UNAP_HIP_SOLVER_PBICGSTAB; no counterpart in the reference. */
int unap_hip_precond_create(unap_hip_ldu A, int type, const unap_hip_amg_params* amg,
                            const double* face_weights, unap_hip_precond* out);
int unap_hip_precond_destroy(unap_hip_precond P);
int unap_hip_precond_apply(unap_hip_precond P, const double* r, double* w, int where);
/* rD() (unapMatrix.hpp:343): reciprocal (preconditioned) diagonal, host copy. */
int unap_hip_precond_rd(unap_hip_precond P, double* rD);
/* Inverse of a dense symmetric positive definite n x n matrix (row-major,
   host arrays), by the blocked Gauss-Jordan inversion on FP64 MFMA that the
   fixed V-cycle uses for its coarsest level.  No counterpart in the
   reference (its coarsest level is solved iteratively, solveCoarsestLevel.cpp). */
int unap_hip_dense_inverse_spd(int64_t n, const double* a, double* inv);
/* Sizes of an AMG preconditioner's hierarchy: level 0 is the matrix itself,
   level l >= 1 the coarse level of agglomeration step l (the reference's
   coarseMatrixLevels(l - 1), lduAgglomeration.hpp:83-92).  Writes up to cap
   entries; *n_levels receives the total count. */
int unap_hip_precond_amg_levels(unap_hip_precond P, int64_t cap, int64_t* n_levels, int64_t* n_cells,
                                int64_t* n_faces);

/* ----------------------------------------------------------------- smoothers
 * Replaces matrix::smoother::smooth(x, A, b, nSweeps) and init()
 * (unapMatrix.hpp:351-373).  The Chebyshev smoother keeps the reference's
 * behaviour: its first call after creation or init() only runs the
 * eigenvalue estimate (eigenDiagPCG, 10 Jacobi-PCG steps that update x). */
int unap_hip_smoother_create(int type, unap_hip_smoother* out);
int unap_hip_smoother_destroy(unap_hip_smoother S);
int unap_hip_smoother_smooth(unap_hip_smoother S, unap_hip_ldu A, double* x, const double* b,
                             int64_t n_sweeps, int where);
int unap_hip_smoother_init(unap_hip_smoother S);
/* eigenDiagPCG (eigenDiagPCG.cpp:11-40): n_pcgs Jacobi-PCG steps on (A, x, b)
   updating x, then the Lanczos/Sturm estimate of the largest eigenvalue. */
int unap_hip_eigen_estimate(unap_hip_ldu A, double* x, const double* b, int64_t n_pcgs,
                            int where, double* lambda_max);

/* ------------------------------------------------------------- agglomeration
 * Replaces lduAgglomeration (lduAgglomeration.hpp:8-97): agglomerate(weights)
 * with SET_maxLevels / SET_nCellsInCoarsestLevel, the per-level maps and the
 * Galerkin coarse matrices (agglomerateMatrix).  Built on the host. */
int unap_hip_aggl_create(unap_hip_ldu A, const double* weights, int64_t max_levels,
                         int64_t n_cells_coarsest, unap_hip_aggl* out);
int unap_hip_aggl_destroy(unap_hip_aggl G);
int unap_hip_aggl_nlevels(unap_hip_aggl G, int64_t* n_levels);
int unap_hip_aggl_level_size(unap_hip_aggl G, int64_t level, int64_t* n_cells, int64_t* n_faces);
/* restrictAddressing / faceRestrictAddressing (lduAgglomeration.hpp:83-92) */
int unap_hip_aggl_restrict_addressing(unap_hip_aggl G, int64_t level, int64_t* out);
int unap_hip_aggl_face_restrict_addressing(unap_hip_aggl G, int64_t level, int64_t* out);
/* coarseMatrixLevels(level) after agglomerateMatrix (lduAgglomerateMatrix.cpp:3-104) */
int unap_hip_aggl_coarse_matrix(unap_hip_aggl G, int64_t level, int64_t* lower_addr,
                                int64_t* upper_addr, double* lower, double* diag, double* upper,
                                int* symmetric);
/* restrictField / prolongField (unapMatrix.hpp:378-438) on the GPU; level is the
   fine level index of the restriction / the coarse level index of the prolongation. */
int unap_hip_aggl_restrict_field(unap_hip_aggl G, int64_t level, const double* fine, double* coarse,
                                 int where);
int unap_hip_aggl_prolong_field(unap_hip_aggl G, int64_t level, const double* coarse, double* fine,
                                int where);

/* ------------------------------------------------------------------ solvers
 * Replaces matrix::solver (unapMatrix.hpp:247-323): PCG (PCG.cpp:35-183),
 * PBiCGStab (PBiCGStab.cpp:45-459), MGSolver (MGSolve.cpp:17-187) and a
 * flexible CG (no reference counterpart; the outer loop of the
 * mixed-precision path).  Defaults as solver::solver() (unapMatrixSolver.cpp:3-10):
 * maxIter 1000, minIter 0, relTol 1e-6, tol 0. */
#define UNAP_HIP_SOLVER_PCG 1
#define UNAP_HIP_SOLVER_PBICGSTAB 2
#define UNAP_HIP_SOLVER_MG 3
#define UNAP_HIP_SOLVER_FCG 4
/* Flexible BiCGStab (no reference counterpart): PBiCGStab's loop
   (PBiCGStab.cpp:45-459), right-preconditioned so that a variable
   preconditioner keeps the x update consistent, with the shadow residual
   restarted (r0 = r) once |r0.r| < 1e-6 |r0| |r|.  Identical to PBiCGStab
   while no restart triggers; meant for FP32 DIC/DILU preconditioners. */
#define UNAP_HIP_SOLVER_FBICGSTAB 5

/* solverPerformance (unapMatrix.hpp:43-174) */
typedef struct {
    double initial_residual;
    double final_residual;
    double previous_residual;
    int64_t n_iterations;
    int32_t converged;
    int32_t singular;
} unap_hip_perf;

int unap_hip_solver_create(int type, unap_hip_solver* out);
int unap_hip_solver_destroy(unap_hip_solver S);
/* PCG(precond) / PBiCGStab(precond) / FCG(precond); NULL = identity */
int unap_hip_solver_set_precond(unap_hip_solver S, unap_hip_precond P);
/* SET_maxIter / SET_minIter / SET_relTol / SET_tolerance */
int unap_hip_solver_set_limits(unap_hip_solver S, int64_t max_iter, int64_t min_iter,
                               double rel_tol, double tol);
/* MGSolver(A, agglomeration, smoothers) with SET_nPreSweeps/nPostSweeps/
   nFinestSweeps; weights NULL -> |upper| / ||upper||_2 (fortranInterface.cpp:383-395).
   The weights (one per face) are copied at the next solve: the array must
   stay valid until then, and later solves need a matrix with that face count. */
int unap_hip_solver_set_mg(unap_hip_solver S, const unap_hip_amg_params* amg, const double* weights);
/* solve(x, A, b): x is the initial guess and the result. */
int unap_hip_solver_solve(unap_hip_solver S, unap_hip_ldu A, double* x, const double* b, int where,
                          unap_hip_perf* perf);
/* Residual history of the last solve: initial residual, then one per iteration
   (at most 2^20 + 2 entries are kept, whatever max_iter is). */
int unap_hip_solver_history(unap_hip_solver S, double* out, int64_t cap, int64_t* len);
/* Device time of the last solve in milliseconds (HIP events around the solve). */
int unap_hip_solver_last_ms(unap_hip_solver S, double* ms);

/* ------------------------------------------------------------- measurement
 * No reference counterpart.  Times `reps` launches of one kernel of the solve
 * path on the library's stream with HIP events, on matrix A's finest level,
 * and returns the mean duration and the kernel's algorithmic HBM bytes per
 * launch (each distinct array read or written counted once; padding, cache
 * re-reads and gathers not counted).  Kernels: */
#define UNAP_HIP_KERNEL_AMUL 1       /* y = A x (lduMatrixSpMV.cpp:16-98) */
#define UNAP_HIP_KERNEL_CHEBY 2      /* one fused Chebyshev sweep (chebySmoother.cpp:105-144) */
#define UNAP_HIP_KERNEL_GS 3         /* one Gauss-Seidel sweep, all levels */
#define UNAP_HIP_KERNEL_DIC 4        /* DIC apply: forward + backward schedule */
#define UNAP_HIP_KERNEL_RESIDUAL 5   /* r = b - A x and sum r^2 (+ finalize) */
#define UNAP_HIP_KERNEL_AXPY_NORM 6  /* x += a p; r -= a w; sum r^2 (+ finalize) */
#define UNAP_HIP_KERNEL_DOT 7        /* sum a b (+ finalize) */
#define UNAP_HIP_KERNEL_COPY 8       /* y = x, a plain copy (achievable-bandwidth yardstick) */
int unap_hip_profile_kernel(unap_hip_ldu A, int kernel, int precision, int64_t reps, double* avg_ms,
                            double* bytes);
/* No counterpart: the setup built on the device (cell orders, sweep
   schedules, SELL-32 layout of each order mode; the agglomeration's pairing
   lists) against the same structures built on the host, for this matrix.
   weights: agglomeration face weights (NULL: |upper|).  *mismatches: the
   number of fields that differ (0: identical). */
int unap_hip_setup_self_check(unap_hip_ldu A, const double* weights, int64_t* mismatches);

/* ------------------------------------------------------- formats and files
 * Host-side construction of matrices from other formats and text files
 * (one domain; no processor interfaces). */
/* coo2ldu after reorderCOO (src/tools/matrixConversion/coo2ldu.cpp:3-115,
   matrixConversion.cpp:38-99): diagonal entries to diag, entries above it to
   faces in row-major order (owner = row < neighbour = column), entries below
   it to lower of the matching face (not read when symmetric).  Any entry
   order (reorderCOO needs the entries grouped by row; a stable row sort comes
   first here); indices from `base` (0 or 1); duplicate entries and a structure that
   is not symmetric are refused (UNAP_HIP_ERR_ARG).  Output arrays may be NULL
   (call once with NULL to learn n_faces, then with arrays of that length;
   diag of length n). */
int unap_hip_coo_to_ldu(int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals,
                        int symmetric, int64_t base, int64_t* n_faces, int64_t* lower_addr, int64_t* upper_addr,
                        double* lower, double* diag, double* upper);
/* A matrix from COO, or from CSR (csr2ldu, coo2ldu.cpp:117-148: row_start of
   length n + 1), by the conversion above. */
int unap_hip_ldu_create_coo(int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals,
                            int symmetric, int64_t base, unap_hip_ldu* out);
int unap_hip_ldu_create_csr(int64_t n, const int64_t* row_start, const int64_t* cols, const double* vals,
                            int symmetric, int64_t base, unap_hip_ldu* out);
/* The OpenFOAM-style text dump: constructLDUMatrixFromOpenFOAM /
   constructVectorFromOpenFOAM (src/tools/readData/readFromOpenFOAM/
   readFromOpenFOAM.cpp:9-192) and printLDUMatrix / printVector
   (src/tools/printData/printUNAP/printUNAP.cpp:8-59, printUNAP.hpp:28-40),
   the same text, at `path` as given (the reference appends "_<rank>.txt"). */
int unap_hip_ldu_read_openfoam(const char* path, unap_hip_ldu* out);
int unap_hip_ldu_write_openfoam(unap_hip_ldu A, const char* path);
int unap_hip_vector_read_openfoam(const char* path, int64_t n, double* out);
int unap_hip_vector_write_openfoam(const char* path, int64_t n, const double* v);
/* The Hypre IJ text format: constructLDUMatrixFromHypre (asymmetric storage)
   and constructVectorFromHypre (src/tools/readData/readFromHypre/
   readFromHypre.cpp:11-254, 454-523), one domain (columns outside the row
   range are refused). */
int unap_hip_ldu_read_hypre(const char* path, unap_hip_ldu* out);
int unap_hip_vector_read_hypre(const char* path, int64_t n, double* out);
/* Reverse Cuthill-McKee order of the cell graph (ReorderingSSM and
   rcmLDU_nowrite, src/tools/rcm/cooRCM.cpp:79-235, 477-637): new_of_old[c] is
   cell c's new label.  Each connected component starts at its unvisited cell
   of least degree (lowest id among equals); neighbours join by ascending
   degree, equal degrees in face order; the sequence is then reversed.  (The
   reference's neighbour sort uses a comparator that is not an ordering, so its
   order is defined only where that does not matter.) */
int unap_hip_rcm_order(int64_t n, int64_t n_faces, const int64_t* lower_addr, const int64_t* upper_addr,
                       int64_t* new_of_old);
/* rcmLDU_rewrite (cooRCM.cpp:639-785) for a given cell order (any
   permutation, e.g. from unap_hip_rcm_order): cells renumbered, each face
   turned so that its owner has the lower new label (its upper and lower
   coefficients swap), faces stably sorted by owner, then neighbour.
   new_face_of_old[f]: old face f's new position, or ~position when it turned
   around (rcmLDU_nowrite's postEdgeOrder).  lower NULL: symmetric (lower is
   upper).  Every output may be NULL. */
int unap_hip_ldu_renumber(int64_t n, int64_t n_faces, const int64_t* lower_addr, const int64_t* upper_addr,
                          const int64_t* new_of_old, const double* lower, const double* diag, const double* upper,
                          int64_t* lower_addr_out, int64_t* upper_addr_out, int64_t* new_face_of_old,
                          double* lower_out, double* diag_out, double* upper_out);

/* ------------------------------------------------------------------ batches
 * Many independent LDU systems solved in one persistent kernel launch: one
 * workgroup runs one system's whole Krylov solve, systems are taken from a
 * work queue largest-first.  System s has cells [cell_off[s], cell_off[s+1])
 * and faces [face_off[s], face_off[s+1]) of the concatenated arrays, with
 * addresses local to the system and faces owner-sorted.  solver[s] is
 * UNAP_HIP_SOLVER_PCG (with DIC) or UNAP_HIP_SOLVER_PBICGSTAB (with DILU);
 * lower may be NULL for symmetric systems (then every system uses upper).
 * The preconditioners run in FP32, the Krylov loop in FP64. */
typedef struct unap_hip_batch_s* unap_hip_batch;
int unap_hip_batch_create(int64_t n_systems, const int64_t* cell_off, const int64_t* face_off,
                          const int64_t* lower_addr, const int64_t* upper_addr, const double* lower,
                          const double* diag, const double* upper, const int32_t* solver,
                          unap_hip_batch* out);
int unap_hip_batch_destroy(unap_hip_batch B);
int unap_hip_batch_solve(unap_hip_batch B, double* x, const double* b, int where, int64_t max_iter,
                         double rel_tol, double tol, unap_hip_perf* perf);
/* Device time of the last batch launch (HIP events), and the launch shape the
   autotuner chose on the first solve (threads per workgroup, workgroups per
   CU, rows a wavefront covers per pass of the vector loops). */
int unap_hip_batch_last_ms(unap_hip_batch B, double* ms);
int unap_hip_batch_config(unap_hip_batch B, int64_t* block_size, int64_t* wg_per_cu,
                          int64_t* rows_per_wavefront);
/* Set the launch shape and skip the autotuner: block_size 256, 512 or 1024,
   wg_per_cu >= 1 with block_size * wg_per_cu <= 2048, rows_per_wavefront 64,
   128 or 256 (results do not depend on it). */
int unap_hip_batch_set_config(unap_hip_batch B, int64_t block_size, int64_t wg_per_cu,
                              int64_t rows_per_wavefront);

#ifdef __cplusplus
}
#endif
#endif
