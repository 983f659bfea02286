/*
 * oracle_capi.h -- flat C interface shared by the CPU oracle (liboracle.so,
 * prefix uo_) and the shim over the compiled reference (oracle/_ref/
 * libunap_ref.so, prefix ur_), so tests drive both with identical calls.
 *
 * TEST INFRASTRUCTURE ONLY (see unap_oracle.hpp).
 *
 * Conventions: labels are int64; `lower == NULL` makes a symmetric matrix
 * (lower and upper share storage, as lduMatrix.cpp:54-57 does for &lower==&upper);
 * perf[6] = {initial, final, nIterations, converged, singular, previous}.
 */
#ifndef UNAP_ORACLE_CAPI_H
#define UNAP_ORACLE_CAPI_H
#include <stdint.h>

#ifndef ORACLE_PREFIX
#define ORACLE_PREFIX uo_
#endif
#define ORACLE_CAT2(a, b) a##b
#define ORACLE_CAT(a, b) ORACLE_CAT2(a, b)
#define OFN(name) ORACLE_CAT(ORACLE_PREFIX, name)

#ifdef __cplusplus
extern "C" {
#endif

void* OFN(ldu_create)(int64_t n, int64_t nf, const int64_t* l, const int64_t* u, const double* lower,
                      const double* diag, const double* upper);
void OFN(ldu_destroy)(void* A);
void OFN(spmv)(void* A, const double* x, double* y);
void OFN(owner_start)(void* A, int64_t* out);   /* n+1 */
void OFN(losort)(void* A, int64_t* out);        /* nf */
void OFN(losort_start)(void* A, int64_t* out);  /* n+1 */

/* type: 0 identity, 1 diagonal, 2 DIC, 3 DILU, 5 one forward Gauss-Seidel
   sweep from zero (oracle only: a non-symmetric linear preconditioner) */
void* OFN(precond_create)(void* A, int type);
void OFN(precond_apply)(void* P, const double* r, double* w, int64_t n);
void OFN(precond_rd)(void* P, double* out, int64_t n);
void OFN(precond_destroy)(void* P);

/* type: 1 Gauss-Seidel, 2 Chebyshev */
void* OFN(smoother_create)(int type);
void OFN(smoother_smooth)(void* S, void* A, double* x, const double* b, int64_t nSweeps);
void OFN(smoother_init)(void* S);
void OFN(smoother_destroy)(void* S);

/* eigenDiagPCG with the diagonal preconditioner of A; x is updated. */
double OFN(eigen_diag_pcg)(void* A, double* x, const double* b, int64_t nPCGs);

void* OFN(aggl_create)(void* A, const double* weights, int64_t maxLevels, int64_t nCoarsest);
int64_t OFN(aggl_nlevels)(void* G);
/* sizes[0] = nCells, sizes[1] = nFaces of coarse level `lev` */
void OFN(aggl_level_sizes)(void* G, int64_t lev, int64_t* sizes);
void OFN(aggl_restrict_addr)(void* G, int64_t lev, int64_t* out);
void OFN(aggl_face_restrict_addr)(void* G, int64_t lev, int64_t* out);
/* Galerkin coefficients of every level (computed once), then read back one. */
void OFN(aggl_build_matrices)(void* G);
void OFN(aggl_coarse_matrix)(void* G, int64_t lev, int64_t* l, int64_t* u, double* lower,
                             double* diag, double* upper, int64_t* symm);
void OFN(aggl_destroy)(void* G);

/* Krylov solve.  solver: 1 PCG, 2 PBiCGStab, 4 FCG, 5 FBiCGStab (oracle only).
   precond: 0 identity, 1 diag, 2 DIC, 3 DILU, 4 AMG (mg[] used), 5 one
   Gauss-Seidel sweep (oracle only).
   mg[] = {smootherType, nCoarsest, nVcycles, nPre, nPost, nFinest} (may be NULL
   unless precond == 4).  history (may be NULL) receives the initial residual
   then one residual per iteration, at most histCap values; the count is
   returned through histLen. */
int OFN(krylov_solve)(void* A, int solver, int precond, const int64_t* mg, int64_t maxIter,
                      int64_t minIter, double relTol, double tol, double* x, const double* b,
                      double* perf, double* history, int64_t histCap, int64_t* histLen);

/* Standalone agglomeration-AMG solve (MGSolver). mg[] as above (nVcycles unused);
   weights NULL -> |upper|/||upper||_2 as mgsolversolve_ does for aggl type 1. */
int OFN(mg_solve)(void* A, const int64_t* mg, const double* weights, int64_t maxIter,
                  int64_t minIter, double relTol, double tol, double* x, const double* b,
                  double* perf, double* history, int64_t histCap, int64_t* histLen);

/* setMatrixCoeffients(diag, upper, lower) (lduMatrix.cpp:508-529): new values,
   same addressing; lower == NULL or == upper keeps one coefficient array. */
void OFN(ldu_set_coeffs)(void* A, const double* diag, const double* upper, const double* lower);

/* A solver whose preconditioner (or MG hierarchy) is built once, from A's
   values at creation, and kept across solves, so that a coefficient update in
   between shows what the reference's objects keep (DIC/DILU/diagonal rD,
   agglomeration maps) and what they recompute (Galerkin matrices on every
   MGSolver::solve).  solver 1, 2, 4 with precond as krylov_solve, or 3 (MG,
   agglomerated with the given face weights, or |upper| / ||upper||_2 when
   weights is null); mg[] as there.  Each run re-arms the
   Chebyshev smoothers (initSmoothers, MG.cpp:54-59) and returns the history
   as krylov_solve does. */
void* OFN(solver_create)(void* A, int solver, int precond, const int64_t* mg, const double* weights);
int OFN(solver_run)(void* S, void* A, int64_t maxIter, int64_t minIter, double relTol, double tol,
                    double* x, const double* b, double* perf, double* history, int64_t histCap,
                    int64_t* histLen);
void OFN(solver_destroy)(void* S);

/* Krylov solves timed with a steady clock, for the bench's CPU-baseline leg:
   seconds[0] = preconditioner setup (including AMG agglomeration),
   seconds[1] = a first solve capped at k1 iterations (cold: the Chebyshev
   smoothers' first call estimates eigenvalues), seconds[2] = a second solve
   from the same x with the same preconditioner capped at k2 iterations.
   x receives the second solve's result, perf its performance. */
int OFN(timed_solve)(void* A, int solver, int precond, const int64_t* mg, int64_t k1, int64_t k2,
                     double relTol, double* x, const double* b, double* perf, double* seconds);

#ifdef __cplusplus
}
#endif
#endif
