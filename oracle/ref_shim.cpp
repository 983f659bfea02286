// ref_shim.cpp -- the oracle's flat C interface (oracle_capi.h, prefix ur_)
// implemented over the reference's own classes.  Built together with the
// reference sources (in place, never copied) into oracle/_ref/libunap_ref.so
// by oracle/Makefile.ref.  TEST INFRASTRUCTURE ONLY: it lets the tests run the
// reference solve path on the same seeded systems as the restatement.
#define ORACLE_PREFIX ur_
#include "oracle_capi.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "MG.hpp"
#include "PBiCGStab.hpp"
#include "PCG.hpp"
#include "chebySmoother.hpp"
#include "eigenDiagPCG.hpp"
#include "lduAgglomeration.hpp"
#include "lduDICPrecond.hpp"
#include "lduDILUPrecond.hpp"
#include "lduDiagPrecond.hpp"
#include "lduGaussSeidelSmoother.hpp"
#include "lduMatrix.hpp"
#include "matrixConversion.hpp"
#include "printUNAP.hpp"
#include "readFromHypre.hpp"
#include "rcmf.h"
#include "readFromOpenFOAM.hpp"
#include "unapMPI.hpp"

using namespace UNAP;

namespace {

void ensureMPI() {
    static bool done = false;
    if (!done) {
        unapMPI::initMPI();
        done = true;
    }
}

scalarField wrap(const double* p, int64_t n) { return scalarField(p, (label)n); }

void fillPerf(const matrix::solverPerformance& p, double* out) {
    if (!out) return;
    out[0] = p.initialResidual();
    out[1] = p.finalResidual();
    out[2] = (double)p.nIterations();
    out[3] = p.converged() ? 1.0 : 0.0;
    out[4] = p.singular() ? 1.0 : 0.0;
    out[5] = p.previousResidual();
}

matrix::preconditioner* makePrecond(const lduMatrix& A, int type) {
    switch (type) {
        case 1: return new lduDiagPrecond(A);
        case 2: return new lduDICPrecond(A);
        case 3: return new lduDILUPrecond(A);
        default: return new matrix::preconditioner;
    }
}

// The agglomeration-AMG preconditioner composed from reference pieces:
// w = 0; MGSolver::solve(w, A, r) with maxIter = minIter = nVcycles, tol = relTol = 0.
struct RefMGPrecond : public matrix::preconditioner {
    const lduMatrix& A;
    lduAgglomeration aggl;
    PtrList<matrix::smoother>* sm;
    MGSolver* mg;
    RefMGPrecond(const lduMatrix& M, int smType, label nCoarsest, label nV, label nPre,
                 label nPost, label nFinest)
        : A(M), aggl(M), sm(nullptr), mg(nullptr) {
        ensureMPI();
        scalarField w(M.upper().size());
        for (label i = 0; i < w.size(); ++i) w[i] = std::fabs(M.upper()[i]);
        aggl.SET_nCellsInCoarsestLevel(nCoarsest);
        aggl.agglomerate(w);
        sm = new PtrList<matrix::smoother>(aggl.size());
        for (label i = 0; i < aggl.size(); ++i) {
            if (smType == 2)
                sm->setLevel(i, *new chebySmoother);
            else
                sm->setLevel(i, *new lduGaussSeidelSmoother);
        }
        mg = new MGSolver(M, aggl, *sm);
        mg->SET_nPreSweeps(nPre);
        mg->SET_nPostSweeps(nPost);
        mg->SET_nFinestSweeps(nFinest);
        mg->SET_maxIter(nV);
        mg->SET_minIter(nV);
        mg->SET_tolerance(0.0);
        mg->SET_relTol(0.0);
    }
    ~RefMGPrecond() {
        delete mg;
        delete sm;
    }
    void precondition(scalarField& w, const scalarField& r) const override {
        w = 0.0;
        mg->solve(w, A, r);
    }
};

// A preconditioner and the Krylov solver that uses it, forwarding the limits
// set on the box to the solver at solve time.
struct OwningKrylov : public matrix::solver {
    std::unique_ptr<matrix::preconditioner> P;
    std::unique_ptr<matrix::solver> S;
    matrix::solverPerformance solve(scalarField& x, const matrix& A,
                                    const scalarField& b) const override {
        S->SET_maxIter(maxIter());
        S->SET_minIter(minIter());
        S->SET_relTol(relTol());
        S->SET_tolerance(tolerance());
        return S->solve(x, A, b);
    }
};

// Run a reference solver repeatedly with maxIter = 1..K to recover the residual
// history the reference prints only in DEBUG builds (it is deterministic).
template <class MakeSolver>
int solveWithHistory(MakeSolver make, const lduMatrix& A, int64_t maxIter, int64_t minIter,
                     double relTol, double tol, double* x, const double* b, double* perf,
                     double* history, int64_t histCap, int64_t* histLen) {
    const int64_t n = A.size();
    scalarField bb = wrap(b, n);
    std::vector<double> x0(x, x + n);
    matrix::solverPerformance finalPerf;
    {
        std::unique_ptr<matrix::solver> S(make());
        S->SET_maxIter(maxIter);
        S->SET_minIter(minIter);
        S->SET_relTol(relTol);
        S->SET_tolerance(tol);
        scalarField xx = wrap(x0.data(), n);
        finalPerf = S->solve(xx, A, bb);
        std::memcpy(x, xx.values(), sizeof(double) * n);
    }
    fillPerf(finalPerf, perf);
    const int64_t nIt = finalPerf.nIterations();
    int64_t len = 0;
    if (history && histCap > 0) {
        history[len++] = finalPerf.initialResidual();
        for (int64_t k = 1; k <= nIt && len < histCap; ++k) {
            std::unique_ptr<matrix::solver> S(make());
            S->SET_maxIter(k);
            S->SET_minIter(minIter);
            S->SET_relTol(relTol);
            S->SET_tolerance(tol);
            scalarField xx = wrap(x0.data(), n);
            matrix::solverPerformance p = S->solve(xx, A, bb);
            history[len++] = p.finalResidual();
        }
    }
    if (histLen) *histLen = history ? len : nIt + 1;
    return 0;
}

}  // namespace

extern "C" {

void* ur_ldu_create(int64_t n, int64_t nf, const int64_t* l, const int64_t* u, const double* lower,
                    const double* diag, const double* upper) {
    labelField la(nf), ua(nf);
    for (int64_t f = 0; f < nf; ++f) {
        la[f] = l[f];
        ua[f] = u[f];
    }
    scalarField dg = wrap(diag, n), up = wrap(upper, nf);
    const label nc = n;
    if (lower == nullptr || lower == upper) return new lduMatrix(nc, la, ua, up, dg, up);
    scalarField lo = wrap(lower, nf);
    return new lduMatrix(nc, la, ua, lo, dg, up);
}

void ur_ldu_destroy(void* A) { delete (lduMatrix*)A; }

void ur_spmv(void* Av, const double* x, double* y) {
    const lduMatrix& A = *(lduMatrix*)Av;
    scalarField xx = wrap(x, A.size()), yy(A.size());
    A.spMV(yy, xx);
    std::memcpy(y, yy.values(), sizeof(double) * A.size());
}

void ur_owner_start(void* Av, int64_t* out) {
    const labelField& v = ((lduMatrix*)Av)->ownerStartAddr();
    for (label i = 0; i < v.size(); ++i) out[i] = v[i];
}
void ur_losort(void* Av, int64_t* out) {
    const labelField& v = ((lduMatrix*)Av)->losortAddr();
    for (label i = 0; i < v.size(); ++i) out[i] = v[i];
}
void ur_losort_start(void* Av, int64_t* out) {
    const labelField& v = ((lduMatrix*)Av)->losortStartAddr();
    for (label i = 0; i < v.size(); ++i) out[i] = v[i];
}

void* ur_precond_create(void* A, int type) { return makePrecond(*(lduMatrix*)A, type); }
void ur_precond_apply(void* P, const double* r, double* w, int64_t n) {
    scalarField rr = wrap(r, n), ww(n);
    ((matrix::preconditioner*)P)->precondition(ww, rr);
    std::memcpy(w, ww.values(), sizeof(double) * n);
}
void ur_precond_rd(void* P, double* out, int64_t n) {
    const scalarField& rd = ((matrix::preconditioner*)P)->rD();
    std::memcpy(out, rd.values(), sizeof(double) * n);
}
void ur_precond_destroy(void* P) { delete (matrix::preconditioner*)P; }

void* ur_smoother_create(int type) {
    if (type == 2) return (matrix::smoother*)new chebySmoother;
    return (matrix::smoother*)new lduGaussSeidelSmoother;
}
void ur_smoother_smooth(void* S, void* Av, double* x, const double* b, int64_t nSweeps) {
    const lduMatrix& A = *(lduMatrix*)Av;
    scalarField xx = wrap(x, A.size()), bb = wrap(b, A.size());
    ((matrix::smoother*)S)->smooth(xx, A, bb, nSweeps);
    std::memcpy(x, xx.values(), sizeof(double) * A.size());
}
void ur_smoother_init(void* S) { ((matrix::smoother*)S)->init(); }
void ur_smoother_destroy(void* S) { delete (matrix::smoother*)S; }

double ur_eigen_diag_pcg(void* Av, double* x, const double* b, int64_t nPCGs) {
    const lduMatrix& A = *(lduMatrix*)Av;
    lduDiagPrecond pre(A);
    scalarField xx = wrap(x, A.size()), bb = wrap(b, A.size());
    eigenDiagPCG e(A, xx, bb, pre, nPCGs);
    std::memcpy(x, xx.values(), sizeof(double) * A.size());
    return e.maxEigenValue();
}

void* ur_aggl_create(void* Av, const double* weights, int64_t maxLevels, int64_t nCoarsest) {
    ensureMPI();
    const lduMatrix& A = *(lduMatrix*)Av;
    lduAgglomeration* g = new lduAgglomeration(A);
    g->SET_maxLevels(maxLevels);
    g->SET_nCellsInCoarsestLevel(nCoarsest);
    scalarField w = wrap(weights, A.upper().size());
    g->agglomerate(w);
    return g;
}
int64_t ur_aggl_nlevels(void* G) { return ((lduAgglomeration*)G)->size(); }
void ur_aggl_level_sizes(void* G, int64_t lev, int64_t* sizes) {
    const lduMatrix& C = ((lduAgglomeration*)G)->coarseMatrixLevels(lev);
    sizes[0] = C.size();
    sizes[1] = C.upperAddr().size();
}
void ur_aggl_restrict_addr(void* G, int64_t lev, int64_t* out) {
    const labelField& v = ((lduAgglomeration*)G)->restrictAddressing(lev);
    for (label i = 0; i < v.size(); ++i) out[i] = v[i];
}
void ur_aggl_face_restrict_addr(void* G, int64_t lev, int64_t* out) {
    const labelField& v = ((lduAgglomeration*)G)->faceRestrictAddressing(lev);
    for (label i = 0; i < v.size(); ++i) out[i] = v[i];
}
void ur_aggl_build_matrices(void* G) {
    matrix::agglomeration& g = *(lduAgglomeration*)G;
    for (label i = 0; i < g.size(); ++i) g.agglomerateMatrix(i);
}
void ur_aggl_coarse_matrix(void* G, int64_t lev, int64_t* l, int64_t* u, double* lower,
                           double* diag, double* upper, int64_t* symm) {
    const lduMatrix& C = ((lduAgglomeration*)G)->coarseMatrixLevels(lev);
    const label nf = C.upperAddr().size();
    for (label f = 0; f < nf; ++f) {
        if (l) l[f] = C.lowerAddr()[f];
        if (u) u[f] = C.upperAddr()[f];
        if (upper) upper[f] = C.upper()[f];
        if (lower) lower[f] = C.lower()[f];
    }
    if (diag)
        for (label i = 0; i < C.size(); ++i) diag[i] = C.diag()[i];
    if (symm) *symm = C.symm() ? 1 : 0;
}
void ur_aggl_destroy(void* G) { delete (lduAgglomeration*)G; }

int ur_krylov_solve(void* Av, int solver, int precond, const int64_t* mg, int64_t maxIter,
                    int64_t minIter, double relTol, double tol, double* x, const double* b,
                    double* perf, double* history, int64_t histCap, int64_t* histLen) {
    const lduMatrix& A = *(lduMatrix*)Av;
    if (solver != 1 && solver != 2) return -2;
    if (precond == 4 && !mg) return -1;
    // each solve gets a fresh preconditioner (and fresh Chebyshev state)
    auto make = [&]() -> matrix::solver* {
        OwningKrylov* o = new OwningKrylov;
        if (precond == 4)
            o->P.reset(new RefMGPrecond(A, (int)mg[0], mg[1], mg[2], mg[3], mg[4], mg[5]));
        else
            o->P.reset(makePrecond(A, precond));
        if (solver == 1)
            o->S.reset(new PCG(*o->P));
        else
            o->S.reset(new PBiCGStab(*o->P));
        return o;
    };
    return solveWithHistory(make, A, maxIter, minIter, relTol, tol, x, b, perf, history, histCap,
                            histLen);
}

int ur_timed_solve(void* Av, int solver, int precond, const int64_t* mg, int64_t k1, int64_t k2,
                   double relTol, double* x, const double* b, double* perf, double* seconds) {
    const lduMatrix& A = *(lduMatrix*)Av;
    if (precond == 4 && !mg) return -1;
    if (solver != 1 && solver != 2) return -2;
    auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<matrix::preconditioner> P;
    if (precond == 4)
        P.reset(new RefMGPrecond(A, (int)mg[0], mg[1], mg[2], mg[3], mg[4], mg[5]));
    else
        P.reset(makePrecond(A, precond));
    std::unique_ptr<matrix::solver> S;
    if (solver == 1)
        S.reset(new PCG(*P));
    else
        S.reset(new PBiCGStab(*P));
    S->SET_relTol(relTol);
    scalarField bb = wrap(b, A.size());
    auto t1 = std::chrono::steady_clock::now();
    // cold: the first preconditioner call runs the smoothers' eigen estimates
    S->SET_maxIter(k1);
    scalarField x1 = wrap(x, A.size());
    S->solve(x1, A, bb);
    auto t2 = std::chrono::steady_clock::now();
    // warm: the same preconditioner again from the same start
    S->SET_maxIter(k2);
    scalarField x2 = wrap(x, A.size());
    matrix::solverPerformance p = S->solve(x2, A, bb);
    auto t3 = std::chrono::steady_clock::now();
    std::memcpy(x, x2.values(), sizeof(double) * A.size());
    fillPerf(p, perf);
    seconds[0] = std::chrono::duration<double>(t1 - t0).count();
    seconds[1] = std::chrono::duration<double>(t2 - t1).count();
    seconds[2] = std::chrono::duration<double>(t3 - t2).count();
    return 0;
}

void ur_ldu_set_coeffs(void* Av, const double* diag, const double* upper, const double* lower) {
    lduMatrix& A = *(lduMatrix*)Av;
    const label n = A.size(), nf = A.upperAddr().size();
    scalarField dg = wrap(diag, n), up = wrap(upper, nf);
    if (lower == nullptr || lower == upper) {
        A.setMatrixCoeffients(dg, up, up);
    } else {
        scalarField lo = wrap(lower, nf);
        A.setMatrixCoeffients(dg, up, lo);
    }
}

// A persistent solver over reference objects (see oracle_capi.h).
struct RefSolverBox {
    int type = 1;
    std::unique_ptr<matrix::preconditioner> P;
    std::unique_ptr<matrix::solver> S;
    // standalone MG
    std::unique_ptr<lduAgglomeration> G;
    PtrList<matrix::smoother>* sm = nullptr;
    MGSolver* mg = nullptr;
    ~RefSolverBox() {
        S.reset();
        delete mg;
        delete sm;
    }
    void arm() const {
        if (mg) mg->initSmoothers();
        if (const RefMGPrecond* p = dynamic_cast<const RefMGPrecond*>(P.get())) p->mg->initSmoothers();
    }
    matrix::solver& solver() const { return mg ? (matrix::solver&)*mg : *S; }
};

void* ur_solver_create(void* Av, int solver, int precond, const int64_t* mg, const double* weights) {
    ensureMPI();
    const lduMatrix& A = *(lduMatrix*)Av;
    RefSolverBox* b = new RefSolverBox;
    b->type = solver;
    if (solver == 3) {
        scalarField w(A.upper().size());
        if (weights) {
            for (label i = 0; i < w.size(); ++i) w[i] = weights[i];
        } else {
            const scalar avg = A.upper().SumSqrt();
            for (label i = 0; i < w.size(); ++i) w[i] = std::fabs(A.upper()[i]) / avg;
        }
        b->G.reset(new lduAgglomeration(A));
        b->G->SET_nCellsInCoarsestLevel(mg[1]);
        b->G->agglomerate(w);
        b->sm = new PtrList<matrix::smoother>(b->G->size());
        for (label i = 0; i < b->G->size(); ++i) {
            if (mg[0] == 2)
                b->sm->setLevel(i, *new chebySmoother);
            else
                b->sm->setLevel(i, *new lduGaussSeidelSmoother);
        }
        b->mg = new MGSolver(A, *b->G, *b->sm);
        b->mg->SET_nPreSweeps(mg[3]);
        b->mg->SET_nPostSweeps(mg[4]);
        b->mg->SET_nFinestSweeps(mg[5]);
        return b;
    }
    if (precond == 4)
        b->P.reset(new RefMGPrecond(A, (int)mg[0], mg[1], mg[2], mg[3], mg[4], mg[5]));
    else
        b->P.reset(makePrecond(A, precond));
    if (solver == 1)
        b->S.reset(new PCG(*b->P));
    else
        b->S.reset(new PBiCGStab(*b->P));
    return b;
}

// The history comes from reruns capped at 1..K iterations with the same
// (persistent) objects, each from the same start and with re-armed smoothers.
int ur_solver_run(void* Sv, void* Av, int64_t maxIter, int64_t minIter, double relTol, double tol,
                  double* x, const double* b, double* perf, double* history, int64_t histCap,
                  int64_t* histLen) {
    const RefSolverBox& B = *(RefSolverBox*)Sv;
    const lduMatrix& A = *(lduMatrix*)Av;
    const int64_t n = A.size();
    scalarField bb = wrap(b, n);
    std::vector<double> x0(x, x + n);
    matrix::solver& S = B.solver();
    auto run = [&](int64_t cap, double* out) {
        B.arm();
        S.SET_maxIter(cap);
        S.SET_minIter(minIter);
        S.SET_relTol(relTol);
        S.SET_tolerance(tol);
        scalarField xx = wrap(x0.data(), n);
        matrix::solverPerformance p = S.solve(xx, A, bb);
        if (out) std::memcpy(out, xx.values(), sizeof(double) * n);
        return p;
    };
    matrix::solverPerformance fin = run(maxIter, x);
    fillPerf(fin, perf);
    int64_t len = 0;
    if (history && histCap > 0) {
        history[len++] = fin.initialResidual();
        for (int64_t k = 1; k <= fin.nIterations() && len < histCap; ++k) history[len++] = run(k, nullptr).finalResidual();
    }
    if (histLen) *histLen = history ? len : fin.nIterations() + 1;
    return 0;
}

void ur_solver_destroy(void* S) { delete (RefSolverBox*)S; }

int ur_mg_solve(void* Av, const int64_t* mg, const double* weights, int64_t maxIter,
                int64_t minIter, double relTol, double tol, double* x, const double* b,
                double* perf, double* history, int64_t histCap, int64_t* histLen) {
    ensureMPI();
    const lduMatrix& A = *(lduMatrix*)Av;
    struct MGBox : public matrix::solver {
        lduAgglomeration aggl;
        PtrList<matrix::smoother>* sm = nullptr;
        MGSolver* mg = nullptr;
        explicit MGBox(const lduMatrix& M) : aggl(M) {}
        ~MGBox() {
            delete mg;
            delete sm;
        }
        matrix::solverPerformance solve(scalarField& x, const matrix& A,
                                        const scalarField& b) const override {
            mg->SET_maxIter(maxIter());
            mg->SET_minIter(minIter());
            mg->SET_relTol(relTol());
            mg->SET_tolerance(tolerance());
            return mg->solve(x, A, b);
        }
    };
    auto make = [&]() -> matrix::solver* {
        MGBox* box = new MGBox(A);
        scalarField w(A.upper().size());
        if (weights) {
            for (label i = 0; i < w.size(); ++i) w[i] = weights[i];
        } else {
            const scalar avg = A.upper().SumSqrt();
            for (label i = 0; i < w.size(); ++i) w[i] = std::fabs(A.upper()[i]) / avg;
        }
        box->aggl.SET_nCellsInCoarsestLevel(mg[1]);
        box->aggl.agglomerate(w);
        box->sm = new PtrList<matrix::smoother>(box->aggl.size());
        for (label i = 0; i < box->aggl.size(); ++i) {
            if (mg[0] == 2)
                box->sm->setLevel(i, *new chebySmoother);
            else
                box->sm->setLevel(i, *new lduGaussSeidelSmoother);
        }
        box->mg = new MGSolver(A, box->aggl, *box->sm);
        box->mg->SET_nPreSweeps(mg[3]);
        box->mg->SET_nPostSweeps(mg[4]);
        box->mg->SET_nFinestSweeps(mg[5]);
        return box;
    };
    return solveWithHistory(make, A, maxIter, minIter, relTol, tol, x, b, perf, history, histCap,
                            histLen);
}

// ---------------------------------------- formats and files (reference tools)
// Arrays of an lduMatrix (lower is upper when symmetric).
void ur_ldu_get(void* Av, int64_t* n, int64_t* nf, int64_t* symm, int64_t* l, int64_t* u, double* lower,
                double* diag, double* upper) {
    const lduMatrix& A = *(lduMatrix*)Av;
    const label m = A.upperAddr().size();
    if (n) *n = A.size();
    if (nf) *nf = m;
    if (symm) *symm = A.symm() ? 1 : 0;
    for (label f = 0; f < m; ++f) {
        if (l) l[f] = A.lowerAddr()[f];
        if (u) u[f] = A.upperAddr()[f];
        if (upper) upper[f] = A.upper()[f];
        if (lower) lower[f] = A.lower()[f];
    }
    if (diag)
        for (label i = 0; i < A.size(); ++i) diag[i] = A.diag()[i];
}

// reorderCOO then coo2ldu (the sequence of coo2ldumatrixcreat_,
// fortranInterface.cpp:74-120) on 0-based copies.
void* ur_coo2ldu(int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals, int symm) {
    ensureMPI();
    std::vector<label> r(rows, rows + nnz), c(cols, cols + nnz);
    std::vector<scalar> v(vals, vals + nnz);
    reorderCOO(v.data(), r.data(), c.data(), n, nnz);
    return &coo2ldu(v.data(), r.data(), c.data(), n, nnz, symm != 0);
}
void* ur_csr2ldu(int64_t n, const int64_t* rowStart, const int64_t* cols, const double* vals, int symm) {
    ensureMPI();
    std::vector<label> rs(rowStart, rowStart + n + 1), c(cols, cols + rowStart[n]);
    return &csr2ldu(vals, rs.data(), c.data(), n, rowStart[n], symm != 0);
}
void ur_reorder_coo(double* val, int64_t* row, int64_t* col, int64_t n, int64_t size) {
    reorderCOO(val, row, col, n, size);
}
void ur_reorder_uface(int64_t* row, int64_t* col, int64_t n, int64_t size, int64_t* newOrder) {
    reorderUFace(row, col, n, size, newOrder);
}
void ur_reorder_lface(int64_t* row, int64_t* col, int64_t n, int64_t size, int64_t* newOrder) {
    reorderLFace(row, col, n, size, newOrder);
}
void ur_reorder_value(double* val, const int64_t* newOrder, int64_t size) { reorderValue(val, newOrder, size); }

// printLDUMatrix / printVector write "<name>_<rank>.txt".
void ur_write_openfoam(void* Av, const char* name) { printLDUMatrix(*(lduMatrix*)Av, name); }
void ur_write_vector_openfoam(const double* v, int64_t n, const char* name) {
    scalarField f(v, n);
    printVector(f, name);
}
void* ur_read_openfoam(const char* path) {
    lduMatrix* A = new lduMatrix;
    constructLDUMatrixFromOpenFOAM(*A, path);
    return A;
}
void ur_read_vector_openfoam(const char* path, int64_t n, double* out) {
    scalarField b(n);
    constructVectorFromOpenFOAM(b, path);
    for (label i = 0; i < n; ++i) out[i] = b[i];
}
// constructLDUMatrixFromHypre also writes "A_diag_<rank>.txt" in the working
// directory (readFromHypre.cpp:226).
void* ur_read_hypre(const char* path) {
    ensureMPI();
    lduMatrix* A = new lduMatrix;
    constructLDUMatrixFromHypre(*A, path);
    return A;
}
void ur_read_vector_hypre(const char* path, int64_t n, double* out) {
    scalarField b(n);
    constructVectorFromHypre(b, path);
    for (label i = 0; i < n; ++i) out[i] = b[i];
}

// rcmLDU_nowrite (cooRCM.cpp:477-637): the new label of every cell and the
// new position of every face (negative-encoded where the face turned).  Its
// cost grows with the square of the face count: small graphs only.
void ur_rcm_ldu(int64_t nf, int64_t n, const int64_t* l, const int64_t* u, int64_t* newOfOld,
                int64_t* newFaceOfOld) {
    rcmLDU_nowrite(nf, n, reinterpret_cast<const label*>(l), reinterpret_cast<const label*>(u),
                   reinterpret_cast<label*>(newOfOld), reinterpret_cast<label*>(newFaceOfOld));
}

}  // extern "C"
