// fortran_shim.cpp -- the reference's Fortran / C entry points
// (src/tools/fortranInterface/fortranInterface.cpp:31-1102) over the C-ABI of
// include/unap_hip.h; declarations and conventions in include/unap_hip_fortran.h.
// Host code only: every solve runs in libunap_hip.so.
#include "unap_hip_fortran.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "unap_hip.h"

namespace {

thread_local int g_status = UNAP_HIP_OK;
bool g_referenceAliasing = false;

struct Fail {
    int code;
    std::string msg;
};
[[noreturn]] void fail(int code, const std::string& m) { throw Fail{code, m}; }
void check(int rc) {
    if (rc != UNAP_HIP_OK) fail(rc, unap_hip_last_error());
}

// Runs f; on failure reports on stderr and records the status (the
// reference's CHECK_POINTER / ERROR_EXIT exit the process instead).
template <class F>
bool run(const char* where, F&& f) {
    try {
        f();
        g_status = UNAP_HIP_OK;
        return true;
    } catch (const Fail& e) {
        g_status = e.code;
        std::fprintf(stderr, "unap_hip: %s: %s\n", where, e.msg.c_str());
    } catch (const std::exception& e) {
        g_status = UNAP_HIP_ERR_INTERNAL;
        std::fprintf(stderr, "unap_hip: %s: %s\n", where, e.what());
    }
    return false;
}

unap_hip_ldu matrixOf(const long* APtrPtr) {
    if (!APtrPtr || *APtrPtr == 0) fail(UNAP_HIP_ERR_ARG, "null matrix handle");
    return reinterpret_cast<unap_hip_ldu>(*APtrPtr);
}
void store(long* ptr, void* obj) {
    if (!ptr) fail(UNAP_HIP_ERR_ARG, "null handle slot");
    *ptr = reinterpret_cast<long>(obj);
}

// Handle of the persistent MG and PBiCGStab solvers: the library solver and
// the settings the reference keeps on its solver object.
struct SwSolver {
    unap_hip_solver S = nullptr;
    unap_hip_precond P = nullptr;
    int64_t maxIter = 1000, minIter = 0;  // solver::solver() (unapMatrixSolver.cpp:3-10)
    double relTol = 1e-6, tol = 0.0;
    // MG only
    unap_hip_amg_params amg;
    std::vector<double> weights;
    bool mgDirty = false;
    ~SwSolver() {
        if (S) unap_hip_solver_destroy(S);
        if (P) unap_hip_precond_destroy(P);
    }
};
SwSolver* solverOf(const long* ptr) {
    if (!ptr || *ptr == 0) fail(UNAP_HIP_ERR_ARG, "null solver handle");
    return reinterpret_cast<SwSolver*>(*ptr);
}

int64_t cellsOf(unap_hip_ldu A) {
    int64_t n = 0;
    check(unap_hip_ldu_info(A, &n, nullptr, nullptr, nullptr, nullptr));
    return n;
}
int64_t facesOf(unap_hip_ldu A) {
    int64_t nf = 0;
    check(unap_hip_ldu_info(A, nullptr, &nf, nullptr, nullptr, nullptr));
    return nf;
}

// One Krylov solve of the one-shot entry points, consuming the matrix.
void krylovOnce(int solverType, double* x, long* APtrPtr, double* b, unap_label* nCells, unap_label* precond,
                double* tol, double* relTol, unap_label* maxIter, unap_label* minIter, unap_label* nIter,
                double* res) {
    if (nIter) *nIter = -1;
    if (res) *res = std::numeric_limits<double>::quiet_NaN();
    unap_hip_ldu A = matrixOf(APtrPtr);
    if (!(x && b && nCells && precond && tol && relTol && maxIter && minIter && nIter && res))
        fail(UNAP_HIP_ERR_ARG, "null argument");
    if (*nCells != cellsOf(A)) fail(UNAP_HIP_ERR_ARG, "nCells does not match the matrix");
    if (*precond < 0 || *precond > 3) fail(UNAP_HIP_ERR_ARG, "precond must be 1 (diagonal), 2 (DIC) or 3 (DILU)");
    unap_hip_precond P = nullptr;
    unap_hip_solver S = nullptr;
    unap_hip_perf perf;
    int rc = UNAP_HIP_OK;
    if (*precond > 0) rc = unap_hip_precond_create(A, (int)*precond, nullptr, nullptr, &P);
    if (rc == UNAP_HIP_OK) rc = unap_hip_solver_create(solverType, &S);
    if (rc == UNAP_HIP_OK) rc = unap_hip_solver_set_precond(S, P);
    if (rc == UNAP_HIP_OK) rc = unap_hip_solver_set_limits(S, *maxIter, *minIter, *relTol, *tol);
    if (rc == UNAP_HIP_OK) rc = unap_hip_solver_solve(S, A, x, b, UNAP_HIP_HOST, &perf);
    const std::string msg = rc == UNAP_HIP_OK ? "" : unap_hip_last_error();
    if (S) unap_hip_solver_destroy(S);
    if (P) unap_hip_precond_destroy(P);
    unap_hip_ldu_destroy(A);  // the one-shot solvers consume the matrix
    *APtrPtr = 0;
    if (rc != UNAP_HIP_OK) fail(rc, msg);
    *nIter = perf.n_iterations;
    *res = perf.final_residual;
}

}  // namespace

extern "C" {

int unap_hip_fortran_last_status(void) { return g_status; }
void unap_hip_fortran_set_reference_aliasing(int on) { g_referenceAliasing = on != 0; }

// fortranInterface.cpp:31-71
void ldumatrixcreat_(long* APtrPtr, unap_label* nCells, unap_label* upperSize, unap_label* lowerAddr,
                     unap_label* upperAddr, double* lower, double* diag, double* upper) {
    run("ldumatrixcreat_", [&] {
        if (!(nCells && upperSize)) fail(UNAP_HIP_ERR_ARG, "null size");
        const bool asym = lower != upper;
        unap_hip_ldu A = nullptr;
        if (asym && g_referenceAliasing) {
            // the reference copies lower into upper through a reference and
            // builds a symmetric matrix of the lower coefficients
            check(unap_hip_ldu_create(*nCells, *upperSize, lowerAddr, upperAddr, nullptr, diag, lower, &A));
        } else {
            check(unap_hip_ldu_create(*nCells, *upperSize, lowerAddr, upperAddr, asym ? lower : nullptr, diag, upper,
                                      &A));
        }
        store(APtrPtr, A);
    });
}

// fortranInterface.cpp:74-120: 1-based COO, sorted, converted
void coo2ldumatrixcreat_(long* APtrPtr, const double* data, const unap_label* rows, const unap_label* cols,
                         const unap_label* nCells, const unap_label* size, const unap_label* symm) {
    run("coo2ldumatrixcreat_", [&] {
        if (!(nCells && size && symm)) fail(UNAP_HIP_ERR_ARG, "null size");
        unap_hip_ldu A = nullptr;
        check(unap_hip_ldu_create_coo(*nCells, *size, rows, cols, data, *symm != 0, 1, &A));
        store(APtrPtr, A);
    });
}

// fortranInterface.cpp:123-144: CSR handed to csr2ldu as given (0-based)
void csr2ldumatrixcreat_(long* APtrPtr, const double* data, const unap_label* compRows, const unap_label* cols,
                         const unap_label* nCells, const unap_label* size, const unap_label* symm) {
    run("csr2ldumatrixcreat_", [&] {
        if (!(nCells && size && symm && compRows)) fail(UNAP_HIP_ERR_ARG, "null argument");
        if (compRows[*nCells] - compRows[0] != *size) fail(UNAP_HIP_ERR_ARG, "size does not match the row starts");
        unap_hip_ldu A = nullptr;
        check(unap_hip_ldu_create_csr(*nCells, compRows, cols, data, *symm != 0, 0, &A));
        store(APtrPtr, A);
    });
}

// fortranInterface.cpp:147-197: one domain, no neighbours
void matrixinterfacescreat_(long* APtrPtr, const unap_label* nNeiProcs, const unap_label*, const unap_label*,
                            const unap_label*, const double*) {
    run("matrixinterfacescreat_", [&] {
        matrixOf(APtrPtr);
        if (nNeiProcs && *nNeiProcs != 0)
            fail(UNAP_HIP_ERR_ARG, "processor interfaces are not supported (one domain on one GPU)");
    });
}

// fortranInterface.cpp:200-268
void pcgsolversolve_(double* x, long* APtrPtr, double* b, unap_label* nCells, unap_label* precond, double* tol,
                     double* relTol, unap_label* maxIter, unap_label* minIter, unap_label* num_iterations,
                     double* final_res_norm) {
    run("pcgsolversolve_", [&] {
        krylovOnce(UNAP_HIP_SOLVER_PCG, x, APtrPtr, b, nCells, precond, tol, relTol, maxIter, minIter, num_iterations,
                   final_res_norm);
    });
}

// fortranInterface.cpp:271-346
void pbicgstabsolversolve_(double* x, long* APtrPtr, double* b, unap_label* nCells, unap_label* precond,
                           double* tol, double* relTol, unap_label* maxIter, unap_label* minIter,
                           unap_label* num_iterations, double* final_res_norm) {
    run("pbicgstabsolversolve_", [&] {
        krylovOnce(UNAP_HIP_SOLVER_PBICGSTAB, x, APtrPtr, b, nCells, precond, tol, relTol, maxIter, minIter,
                   num_iterations, final_res_norm);
    });
}

// fortranInterface.cpp:349-464: MGSolver with nPreSweeps 0, the default
// level limits, weights |upper| / ||upper||_2 (aggl 1) or faceArea (aggl 2)
void mgsolversolve_(double* x, long* APtrPtr, double* b, unap_label* nCells, unap_label* aggl, unap_label* smoother,
                    double* tol, double* relTol, unap_label* maxIter, unap_label* minIter,
                    unap_label* num_iterations, double* final_res_norm, double* faceArea) {
    run("mgsolversolve_", [&] {
        if (num_iterations) *num_iterations = -1;
        if (final_res_norm) *final_res_norm = std::numeric_limits<double>::quiet_NaN();
        unap_hip_ldu A = matrixOf(APtrPtr);
        if (!(x && b && nCells && aggl && smoother && tol && relTol && maxIter && minIter && num_iterations &&
              final_res_norm))
            fail(UNAP_HIP_ERR_ARG, "null argument");
        if (*nCells != cellsOf(A)) fail(UNAP_HIP_ERR_ARG, "nCells does not match the matrix");
        if (*aggl != 1 && *aggl != 2) fail(UNAP_HIP_ERR_ARG, "aggl must be 1 (|upper|) or 2 (face area)");
        if (*aggl == 2 && !faceArea) fail(UNAP_HIP_ERR_ARG, "aggl 2 needs faceArea");
        unap_hip_amg_params prm;
        check(unap_hip_amg_params_default(&prm));
        prm.smoother = *smoother;
        prm.n_pre_sweeps = 0;
        unap_hip_solver S = nullptr;
        unap_hip_perf perf;
        int rc = unap_hip_solver_create(UNAP_HIP_SOLVER_MG, &S);
        if (rc == UNAP_HIP_OK) rc = unap_hip_solver_set_mg(S, &prm, *aggl == 2 ? faceArea : nullptr);
        if (rc == UNAP_HIP_OK) rc = unap_hip_solver_set_limits(S, *maxIter, *minIter, *relTol, *tol);
        if (rc == UNAP_HIP_OK) rc = unap_hip_solver_solve(S, A, x, b, UNAP_HIP_HOST, &perf);
        const std::string msg = rc == UNAP_HIP_OK ? "" : unap_hip_last_error();
        if (S) unap_hip_solver_destroy(S);
        unap_hip_ldu_destroy(A);
        *APtrPtr = 0;
        if (rc != UNAP_HIP_OK) fail(rc, msg);
        *num_iterations = perf.n_iterations;
        *final_res_norm = perf.final_residual;
    });
}

// reorderCOO (matrixConversion.cpp:38-99): with the entries grouped by row,
// each row's entries placed in ascending column order (an entry's place in
// its row is the number of entries of that row with a smaller column).
// reordercoo_ (fortranInterface.cpp:473-489) first makes the indices 0-based
// when the first row index is not 0.
void reordercoo_(double* val, unap_label* row, unap_label* col, unap_label* nCells, unap_label* size) {
    run("reordercoo_", [&] {
        const int64_t n = *nCells, m = *size;
        if (m > 0 && row[0] != 0)
            for (int64_t i = 0; i < m; ++i) {
                row[i] -= 1;
                col[i] -= 1;
            }
        std::vector<int64_t> off(n + 1, 0), pos(m, 0);
        for (int64_t i = 0; i < m; ++i) {
            if (row[i] < 0 || row[i] >= n) fail(UNAP_HIP_ERR_ARG, "row index out of range");
            off[row[i] + 1]++;
        }
        for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
        for (int64_t i = 0; i < m; ++i)
            for (int64_t k = off[row[i]]; k < off[row[i] + 1]; ++k)
                if (col[i] > col[k]) pos[i]++;
        std::vector<int64_t> r2(m), c2(m);
        std::vector<double> v2(m);
        for (int64_t i = 0; i < m; ++i) {
            const int64_t p = off[row[i]] + pos[i];
            v2[p] = val[i];
            r2[p] = row[i];
            c2[p] = col[i];
        }
        for (int64_t i = 0; i < m; ++i) {
            val[i] = v2[i];
            row[i] = r2[i];
            col[i] = c2[i];
        }
    });
}

// reorderUFace (matrixConversion.cpp:117-172) after making the indices
// 0-based (fortranInterface.cpp:492-507): the same placement as reorderCOO,
// recorded in newOrder instead of applied to values.
void reorderuface__(unap_label* row, unap_label* col, unap_label* nCells, unap_label* size, unap_label* newOrder) {
    run("reorderuface__", [&] {
        const int64_t n = *nCells, m = *size;
        for (int64_t i = 0; i < m; ++i) {
            row[i] -= 1;
            col[i] -= 1;
            if (row[i] < 0 || row[i] >= n) fail(UNAP_HIP_ERR_ARG, "row index out of range");
        }
        std::vector<int64_t> off(n + 1, 0), pos(m, 0);
        for (int64_t i = 0; i < m; ++i) off[row[i] + 1]++;
        for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
        for (int64_t i = 0; i < m; ++i)
            for (int64_t k = off[row[i]]; k < off[row[i] + 1]; ++k)
                if (col[i] > col[k]) pos[i]++;
        std::vector<int64_t> r2(m), c2(m);
        for (int64_t i = 0; i < m; ++i) {
            const int64_t p = off[row[i]] + pos[i];
            newOrder[i] = p;
            r2[p] = row[i];
            c2[p] = col[i];
        }
        for (int64_t i = 0; i < m; ++i) {
            row[i] = r2[i];
            col[i] = c2[i];
        }
    });
}

// reorderLFace (matrixConversion.cpp:175-215) after making the indices
// 0-based (fortranInterface.cpp:510-525): entries grouped by column, in
// input order within a column.
void reorderlface__(unap_label* row, unap_label* col, unap_label* nCells, unap_label* size, unap_label* newOrder) {
    run("reorderlface__", [&] {
        const int64_t n = *nCells, m = *size;
        for (int64_t i = 0; i < m; ++i) {
            row[i] -= 1;
            col[i] -= 1;
            if (col[i] < 0 || col[i] >= n) fail(UNAP_HIP_ERR_ARG, "column index out of range");
        }
        std::vector<int64_t> off(n + 1, 0), cnt(n, 0);
        for (int64_t i = 0; i < m; ++i) off[col[i] + 1]++;
        for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
        std::vector<int64_t> r2(m), c2(m);
        for (int64_t i = 0; i < m; ++i) {
            const int64_t p = off[col[i]] + cnt[col[i]]++;
            newOrder[i] = p;
            r2[p] = row[i];
            c2[p] = col[i];
        }
        for (int64_t i = 0; i < m; ++i) {
            row[i] = r2[i];
            col[i] = c2[i];
        }
    });
}

// reorderValue (matrixConversion.cpp:102-114): val[newOrder[i]] = old val[i].
void reordervalue__(double* val, unap_label* newOrder, unap_label* size) {
    run("reordervalue__", [&] {
        const std::vector<double> old(val, val + *size);
        for (int64_t i = 0; i < *size; ++i) val[newOrder[i]] = old[i];
    });
}

// fortranInterface.cpp:537-573: addressing now (0-based, as given),
// coefficients zero until fill_sw_matrix_coefficients__
void contruct_sw_matrix__(long* APtrPtr, const unap_label* nCells, const unap_label* rows, const unap_label* cols,
                          const unap_label* size) {
    run("contruct_sw_matrix__", [&] {
        if (!(nCells && size)) fail(UNAP_HIP_ERR_ARG, "null size");
        const std::vector<double> d(*nCells, 0.0), up(*size, 0.0);
        unap_hip_ldu A = nullptr;
        check(unap_hip_ldu_create(*nCells, *size, rows, cols, nullptr, d.data(), up.data(), &A));
        store(APtrPtr, A);
    });
}

// fortranInterface.cpp:576-615: one domain, no neighbours
void contruct_sw_matrix_interfaces__(long* APtrPtr, const unap_label* nNeiProcs, const unap_label*,
                                     const unap_label*, const unap_label*) {
    run("contruct_sw_matrix_interfaces__", [&] {
        matrixOf(APtrPtr);
        if (nNeiProcs && *nNeiProcs != 0)
            fail(UNAP_HIP_ERR_ARG, "processor interfaces are not supported (one domain on one GPU)");
    });
}

// fortranInterface.cpp:629-672: symmetric when upper and lower are the same array
void fill_sw_matrix_coefficients__(long* APtrPtr, const double* diag, const double* upper, const double* lower) {
    run("fill_sw_matrix_coefficients__", [&] {
        check(unap_hip_ldu_set_coeffs(matrixOf(APtrPtr), diag, upper, upper == lower ? nullptr : lower));
    });
}

// fortranInterface.cpp:675-699: no interfaces, nothing to fill
void fill_sw_matrix_interfaces_coefficients__(long* APtrPtr, const unap_label*, const double*) {
    run("fill_sw_matrix_interfaces_coefficients__", [&] { matrixOf(APtrPtr); });
}

// fortranInterface.cpp:703-770
void contruct_solver_mg__(long* mgPtrPtr, long* APtrPtr, long* AgglPtr, const double* weights,
                          const unap_label* smootherType, const unap_label* maxLevels,
                          const unap_label* nCellsCoarsest) {
    run("contruct_solver_mg__", [&] {
        unap_hip_ldu A = matrixOf(APtrPtr);
        if (!(weights && smootherType && maxLevels && nCellsCoarsest)) fail(UNAP_HIP_ERR_ARG, "null argument");
        if (*smootherType != 1 && *smootherType != 2)
            fail(UNAP_HIP_ERR_ARG, "smoother must be 1 (Gauss-Seidel) or 2 (Chebyshev)");
        SwSolver* s = new SwSolver;
        s->weights.assign(weights, weights + facesOf(A));
        check(unap_hip_amg_params_default(&s->amg));
        s->amg.smoother = *smootherType;
        s->amg.max_levels = *maxLevels;
        s->amg.n_cells_coarsest = *nCellsCoarsest;
        s->mgDirty = true;
        if (unap_hip_solver_create(UNAP_HIP_SOLVER_MG, &s->S) != UNAP_HIP_OK) {
            const std::string m = unap_hip_last_error();
            delete s;
            fail(UNAP_HIP_ERR_HIP, m);
        }
        store(mgPtrPtr, s);
        if (AgglPtr) store(AgglPtr, s);  // the agglomeration lives inside the solver
    });
}

void sw_solver_mg_set_maxiter__(long* p, const unap_label* v) {
    run("sw_solver_mg_set_maxiter__", [&] { solverOf(p)->maxIter = *v; });
}
void sw_solver_mg_set_miniter__(long* p, const unap_label* v) {
    run("sw_solver_mg_set_miniter__", [&] { solverOf(p)->minIter = *v; });
}
void sw_solver_mg_set_tol__(long* p, const double* v) {
    run("sw_solver_mg_set_tol__", [&] { solverOf(p)->tol = *v; });
}
void sw_solver_mg_set_reltol__(long* p, const double* v) {
    run("sw_solver_mg_set_reltol__", [&] { solverOf(p)->relTol = *v; });
}
void sw_solver_mg_set_npresweeps__(long* p, const unap_label* v) {
    run("sw_solver_mg_set_npresweeps__", [&] {
        SwSolver* s = solverOf(p);
        s->amg.n_pre_sweeps = *v;
        s->mgDirty = true;
    });
}
void sw_solver_mg_set_npostsweeps__(long* p, const unap_label* v) {
    run("sw_solver_mg_set_npostsweeps__", [&] {
        SwSolver* s = solverOf(p);
        s->amg.n_post_sweeps = *v;
        s->mgDirty = true;
    });
}
void sw_solver_mg_set_nfinestsweeps__(long* p, const unap_label* v) {
    run("sw_solver_mg_set_nfinestsweeps__", [&] {
        SwSolver* s = solverOf(p);
        s->amg.n_finest_sweeps = *v;
        s->mgDirty = true;
    });
}

// fortranInterface.cpp:870-922 (initSmoothers is part of every library solve)
void sw_solve_mg__(long* mgPtrPtr, long* APtrPtr, double* x, double* b, double* res_norm) {
    run("sw_solve_mg__", [&] {
        if (res_norm) *res_norm = std::numeric_limits<double>::quiet_NaN();
        SwSolver* s = solverOf(mgPtrPtr);
        unap_hip_ldu A = matrixOf(APtrPtr);
        if (!(x && b && res_norm)) fail(UNAP_HIP_ERR_ARG, "null argument");
        if (s->mgDirty) {
            if ((int64_t)s->weights.size() != facesOf(A)) fail(UNAP_HIP_ERR_ARG, "weights were given for another matrix");
            check(unap_hip_solver_set_mg(s->S, &s->amg, s->weights.data()));
            s->mgDirty = false;
        }
        check(unap_hip_solver_set_limits(s->S, s->maxIter, s->minIter, s->relTol, s->tol));
        unap_hip_perf perf;
        check(unap_hip_solver_solve(s->S, A, x, b, UNAP_HIP_HOST, &perf));
        *res_norm = perf.final_residual;
    });
}

// fortranInterface.cpp:926-933: no preconditioner until set_precond
void contruct_solver_pbicgstab__(long* solverPtrPtr) {
    run("contruct_solver_pbicgstab__", [&] {
        SwSolver* s = new SwSolver;
        if (unap_hip_solver_create(UNAP_HIP_SOLVER_PBICGSTAB, &s->S) != UNAP_HIP_OK) {
            const std::string m = unap_hip_last_error();
            delete s;
            fail(UNAP_HIP_ERR_HIP, m);
        }
        store(solverPtrPtr, s);
    });
}
void sw_solver_pbicgstab_set_maxiter__(long* p, const unap_label* v) {
    run("sw_solver_pbicgstab_set_maxiter__", [&] { solverOf(p)->maxIter = *v; });
}
void sw_solver_pbicgstab_set_miniter__(long* p, const unap_label* v) {
    run("sw_solver_pbicgstab_set_miniter__", [&] { solverOf(p)->minIter = *v; });
}
void sw_solver_pbicgstab_set_tol__(long* p, const double* v) {
    run("sw_solver_pbicgstab_set_tol__", [&] { solverOf(p)->tol = *v; });
}
void sw_solver_pbicgstab_set_reltol__(long* p, const double* v) {
    run("sw_solver_pbicgstab_set_reltol__", [&] { solverOf(p)->relTol = *v; });
}

// fortranInterface.cpp:972-1004: the preconditioner is built now, from the
// matrix's current coefficients
void sw_solver_pbicgstab_set_precond__(long* solverPtrPtr, long* APtrPtr, const unap_label* precondType) {
    run("sw_solver_pbicgstab_set_precond__", [&] {
        SwSolver* s = solverOf(solverPtrPtr);
        unap_hip_ldu A = matrixOf(APtrPtr);
        if (!precondType || *precondType < 1 || *precondType > 3)
            fail(UNAP_HIP_ERR_ARG, "precondition type must be 1 (diagonal), 2 (DIC) or 3 (DILU)");
        unap_hip_precond P = nullptr;
        check(unap_hip_precond_create(A, (int)*precondType, nullptr, nullptr, &P));
        check(unap_hip_solver_set_precond(s->S, P));
        if (s->P) unap_hip_precond_destroy(s->P);
        s->P = P;
    });
}

// fortranInterface.cpp:1007-1058
void sw_solve_pbicgstab__(long* solverPtrPtr, long* APtrPtr, double* x, double* b, double* res_norm) {
    run("sw_solve_pbicgstab__", [&] {
        if (res_norm) *res_norm = std::numeric_limits<double>::quiet_NaN();
        SwSolver* s = solverOf(solverPtrPtr);
        unap_hip_ldu A = matrixOf(APtrPtr);
        if (!(x && b && res_norm)) fail(UNAP_HIP_ERR_ARG, "null argument");
        check(unap_hip_solver_set_limits(s->S, s->maxIter, s->minIter, s->relTol, s->tol));
        unap_hip_perf perf;
        check(unap_hip_solver_solve(s->S, A, x, b, UNAP_HIP_HOST, &perf));
        *res_norm = perf.final_residual;
    });
}

// fortranInterface.cpp:1061-1102
void sw_matrix_destroy__(long* APtrPtr) {
    run("sw_matrix_destroy__", [&] {
        check(unap_hip_ldu_destroy(matrixOf(APtrPtr)));
        *APtrPtr = 0;
    });
}
void sw_solver_destroy_mg__(long* p) {
    run("sw_solver_destroy_mg__", [&] {
        delete solverOf(p);
        *p = 0;
    });
}
void sw_solver_destroy_pbicgstab__(long* p) {
    run("sw_solver_destroy_pbicgstab__", [&] {
        delete solverOf(p);
        *p = 0;
    });
}

}  // extern "C"
