// oracle_capi.cpp -- flat C entry points of the CPU oracle (test
// infrastructure only; see unap_oracle.hpp).
#define ORACLE_PREFIX uo_
#include "oracle_capi.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "unap_oracle.hpp"

using namespace unap_oracle;

namespace {

void fillPerf(const Perf& p, double* out) {
    if (!out) return;
    out[0] = p.initial;
    out[1] = p.final_;
    out[2] = (double)p.nIter;
    out[3] = p.converged ? 1.0 : 0.0;
    out[4] = p.singular ? 1.0 : 0.0;
    out[5] = p.previous;
}

void copyHistory(const std::vector<double>& h, double* out, int64_t cap, int64_t* len) {
    if (len) *len = (int64_t)h.size();
    if (!out) return;
    const int64_t m = (int64_t)h.size() < cap ? (int64_t)h.size() : cap;
    std::memcpy(out, h.data(), sizeof(double) * m);
}

Precond* makePrecond(const Ldu& A, int type) {
    switch (type) {
        case 1: return new DiagPrecond(A);
        case 2: return new DICPrecond(A);
        case 3: return new DILUPrecond(A);
        case 5: return new GSPrecond(A);
        default: return new Precond;
    }
}

struct SmootherBox {
    int type;
    std::unique_ptr<Smoother> s;
};

}  // namespace

extern "C" {

void* uo_ldu_create(int64_t n, int64_t nf, const int64_t* l, const int64_t* u, const double* lower,
                    const double* diag, const double* upper) {
    Ldu* A = new Ldu;
    A->n = n;
    A->l.assign(l, l + nf);
    A->u.assign(u, u + nf);
    A->d.assign(diag, diag + n);
    A->up.assign(upper, upper + nf);
    A->symm = (lower == nullptr || lower == upper);
    if (!A->symm) A->lo.assign(lower, lower + nf);
    return A;
}

void uo_ldu_destroy(void* A) { delete (Ldu*)A; }

void uo_spmv(void* A, const double* x, double* y) { ((Ldu*)A)->spmv(y, x); }

void uo_owner_start(void* A, int64_t* out) {
    const auto& v = ((Ldu*)A)->ownerStart();
    std::memcpy(out, v.data(), sizeof(int64_t) * v.size());
}
void uo_losort(void* A, int64_t* out) {
    const auto& v = ((Ldu*)A)->losort();
    if (!v.empty()) std::memcpy(out, v.data(), sizeof(int64_t) * v.size());
}
void uo_losort_start(void* A, int64_t* out) {
    const auto& v = ((Ldu*)A)->losortStart();
    std::memcpy(out, v.data(), sizeof(int64_t) * v.size());
}

struct PrecBox {
    std::unique_ptr<Precond> p;
};

void* uo_precond_create(void* A, int type) { return new PrecBox{std::unique_ptr<Precond>(makePrecond(*(Ldu*)A, type))}; }
void uo_precond_apply(void* P, const double* r, double* w, int64_t n) { ((PrecBox*)P)->p->apply(w, r, n); }
void uo_precond_rd(void* P, double* out, int64_t n) {
    const double* rd = ((PrecBox*)P)->p->rD();
    if (rd) std::memcpy(out, rd, sizeof(double) * n);
}
void uo_precond_destroy(void* P) { delete (PrecBox*)P; }

void* uo_smoother_create(int type) {
    SmootherBox* b = new SmootherBox;
    b->type = type;
    if (type == 2)
        b->s.reset(new ChebySmoother);
    else
        b->s.reset(new GSSmoother);
    return b;
}
void uo_smoother_smooth(void* S, void* A, double* x, const double* b, int64_t nSweeps) {
    ((SmootherBox*)S)->s->smooth(x, *(Ldu*)A, b, nSweeps);
}
void uo_smoother_init(void* S) { ((SmootherBox*)S)->s->init(); }
void uo_smoother_destroy(void* S) { delete (SmootherBox*)S; }

double uo_eigen_diag_pcg(void* A, double* x, const double* b, int64_t nPCGs) {
    DiagPrecond pre(*(Ldu*)A);
    return eigenDiagPCG(*(Ldu*)A, x, b, pre.rd.data(), nPCGs);
}

void* uo_aggl_create(void* A, const double* weights, int64_t maxLevels, int64_t nCoarsest) {
    Agglomeration* g = new Agglomeration(*(Ldu*)A);
    g->maxLevels = maxLevels;
    g->nCoarsest = nCoarsest;
    g->agglomerate(weights);
    return g;
}
int64_t uo_aggl_nlevels(void* G) { return ((Agglomeration*)G)->size(); }
void uo_aggl_level_sizes(void* G, int64_t lev, int64_t* sizes) {
    const Ldu& C = ((Agglomeration*)G)->coarse[lev];
    sizes[0] = C.n;
    sizes[1] = C.nFaces();
}
void uo_aggl_restrict_addr(void* G, int64_t lev, int64_t* out) {
    const auto& v = ((Agglomeration*)G)->restrictAddr[lev];
    std::memcpy(out, v.data(), sizeof(int64_t) * v.size());
}
void uo_aggl_face_restrict_addr(void* G, int64_t lev, int64_t* out) {
    const auto& v = ((Agglomeration*)G)->faceRestrictAddr[lev];
    if (!v.empty()) std::memcpy(out, v.data(), sizeof(int64_t) * v.size());
}
void uo_aggl_build_matrices(void* G) { ((Agglomeration*)G)->agglomerateAll(); }
void uo_aggl_coarse_matrix(void* G, int64_t lev, int64_t* l, int64_t* u, double* lower,
                           double* diag, double* upper, int64_t* symm) {
    const Ldu& C = ((Agglomeration*)G)->coarse[lev];
    const int64_t nf = C.nFaces();
    if (l) std::memcpy(l, C.l.data(), sizeof(int64_t) * nf);
    if (u) std::memcpy(u, C.u.data(), sizeof(int64_t) * nf);
    if (diag) std::memcpy(diag, C.d.data(), sizeof(double) * C.n);
    if (upper) std::memcpy(upper, C.upper(), sizeof(double) * nf);
    if (lower) std::memcpy(lower, C.lower(), sizeof(double) * nf);
    if (symm) *symm = C.symm ? 1 : 0;
}
void uo_aggl_destroy(void* G) { delete (Agglomeration*)G; }

int uo_krylov_solve(void* Av, int solver, int precond, const int64_t* mg, int64_t maxIter,
                    int64_t minIter, double relTol, double tol, double* x, const double* b,
                    double* perf, double* history, int64_t histCap, int64_t* histLen) {
    const Ldu& A = *(Ldu*)Av;
    std::unique_ptr<Precond> P;
    if (precond == 4) {
        if (!mg) return -1;
        P.reset(new MGPrecond(A, (int)mg[0], mg[1], mg[2], mg[3], mg[4], mg[5]));
    } else {
        P.reset(makePrecond(A, precond));
    }
    std::unique_ptr<Solver> S;
    if (solver == 1)
        S.reset(new PCG(*P));
    else if (solver == 2)
        S.reset(new PBiCGStab(*P));
    else if (solver == 4)
        S.reset(new FCG(*P));
    else if (solver == 5)
        S.reset(new FBiCGStab(*P));
    else
        return -2;
    S->maxIter = maxIter;
    S->minIter = minIter;
    S->relTol = relTol;
    S->tol = tol;
    std::vector<double> h;
    S->history = &h;
    Perf p = S->solve(x, A, b);
    fillPerf(p, perf);
    copyHistory(h, history, histCap, histLen);
    return 0;
}

int uo_mg_solve(void* Av, const int64_t* mg, const double* weights, int64_t maxIter,
                int64_t minIter, double relTol, double tol, double* x, const double* b,
                double* perf, double* history, int64_t histCap, int64_t* histLen) {
    const Ldu& A = *(Ldu*)Av;
    std::vector<double> w(A.nFaces());
    if (weights) {
        for (int64_t f = 0; f < A.nFaces(); ++f) w[f] = weights[f];
    } else {
        // |upper| / ||upper||_2 (fortranInterface.cpp:383-395)
        const double avg = std::sqrt(sumSqr(A.upper(), A.nFaces()));
        for (int64_t f = 0; f < A.nFaces(); ++f) w[f] = std::fabs(A.upper()[f]) / avg;
    }
    Agglomeration G(A);
    G.nCoarsest = mg[1];
    G.agglomerate(w.data());
    G.agglomerateAll();
    std::vector<std::unique_ptr<Smoother>> owned;
    std::vector<const Smoother*> sm;
    for (int64_t i = 0; i < G.size(); ++i) {
        if (mg[0] == 2)
            owned.emplace_back(new ChebySmoother);
        else
            owned.emplace_back(new GSSmoother);
        sm.push_back(owned.back().get());
    }
    MGSolver S(A, G, sm);
    S.nPre = mg[3];
    S.nPost = mg[4];
    S.nFinest = mg[5];
    S.maxIter = maxIter;
    S.minIter = minIter;
    S.relTol = relTol;
    S.tol = tol;
    std::vector<double> h;
    S.history = &h;
    Perf p = S.solve(x, A, b);
    fillPerf(p, perf);
    copyHistory(h, history, histCap, histLen);
    return 0;
}

void uo_ldu_set_coeffs(void* Av, const double* diag, const double* upper, const double* lower) {
    Ldu& A = *(Ldu*)Av;
    const int64_t nf = A.nFaces();
    A.d.assign(diag, diag + A.n);
    A.up.assign(upper, upper + nf);
    A.symm = (lower == nullptr || lower == upper);
    A.lo.clear();
    if (!A.symm) A.lo.assign(lower, lower + nf);
}

// A persistent solver: preconditioner or MG hierarchy built once.
struct SolverBox {
    int type = 1;
    std::unique_ptr<Precond> P;
    std::unique_ptr<Solver> S;
    // standalone MG
    std::unique_ptr<Agglomeration> G;
    std::vector<std::unique_ptr<Smoother>> sm;
};

void* uo_solver_create(void* Av, int solver, int precond, const int64_t* mg, const double* weights) {
    const Ldu& A = *(Ldu*)Av;
    std::unique_ptr<SolverBox> b(new SolverBox);
    b->type = solver;
    if (solver == 3) {
        if (!mg) return nullptr;
        std::vector<double> w(A.nFaces());
        if (weights) {
            for (int64_t f = 0; f < A.nFaces(); ++f) w[f] = weights[f];
        } else {
            const double avg = std::sqrt(sumSqr(A.upper(), A.nFaces()));
            for (int64_t f = 0; f < A.nFaces(); ++f) w[f] = std::fabs(A.upper()[f]) / avg;
        }
        b->G.reset(new Agglomeration(A));
        b->G->nCoarsest = mg[1];
        b->G->agglomerate(w.data());
        std::vector<const Smoother*> sp;
        for (int64_t i = 0; i < b->G->size(); ++i) {
            b->sm.emplace_back(mg[0] == 2 ? (Smoother*)new ChebySmoother : (Smoother*)new GSSmoother);
            sp.push_back(b->sm.back().get());
        }
        MGSolver* M = new MGSolver(A, *b->G, sp);
        M->nPre = mg[3];
        M->nPost = mg[4];
        M->nFinest = mg[5];
        b->S.reset(M);
        return b.release();
    }
    if (precond == 4) {
        if (!mg) return nullptr;
        b->P.reset(new MGPrecond(A, (int)mg[0], mg[1], mg[2], mg[3], mg[4], mg[5]));
    } else {
        b->P.reset(makePrecond(A, precond));
    }
    if (solver == 1)
        b->S.reset(new PCG(*b->P));
    else if (solver == 2)
        b->S.reset(new PBiCGStab(*b->P));
    else if (solver == 4)
        b->S.reset(new FCG(*b->P));
    else if (solver == 5)
        b->S.reset(new FBiCGStab(*b->P));
    else
        return nullptr;
    return b.release();
}

int uo_solver_run(void* Sv, void* Av, int64_t maxIter, int64_t minIter, double relTol, double tol,
                  double* x, const double* b, double* perf, double* history, int64_t histCap,
                  int64_t* histLen) {
    SolverBox& B = *(SolverBox*)Sv;
    const Ldu& A = *(Ldu*)Av;
    for (auto& s : B.sm) s->init();
    if (MGPrecond* mp = dynamic_cast<MGPrecond*>(B.P.get())) mp->init();
    B.S->maxIter = maxIter;
    B.S->minIter = minIter;
    B.S->relTol = relTol;
    B.S->tol = tol;
    std::vector<double> h;
    B.S->history = &h;
    Perf p = B.S->solve(x, A, b);
    B.S->history = nullptr;
    fillPerf(p, perf);
    copyHistory(h, history, histCap, histLen);
    return 0;
}

void uo_solver_destroy(void* S) { delete (SolverBox*)S; }

int uo_timed_solve(void* Av, int solver, int precond, const int64_t* mg, int64_t k1, int64_t k2,
                   double relTol, double* x, const double* b, double* perf, double* seconds) {
    const Ldu& A = *(Ldu*)Av;
    auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<Precond> P;
    if (precond == 4)
        P.reset(new MGPrecond(A, (int)mg[0], mg[1], mg[2], mg[3], mg[4], mg[5]));
    else
        P.reset(makePrecond(A, precond));
    std::unique_ptr<Solver> S;
    if (solver == 1)
        S.reset(new PCG(*P));
    else if (solver == 2)
        S.reset(new PBiCGStab(*P));
    else
        S.reset(new FCG(*P));
    S->relTol = relTol;
    std::vector<double> x1(x, x + A.n), x2(x, x + A.n);
    auto t1 = std::chrono::steady_clock::now();
    S->maxIter = k1;
    S->solve(x1.data(), A, b);
    auto t2 = std::chrono::steady_clock::now();
    S->maxIter = k2;
    Perf p = S->solve(x2.data(), A, b);
    auto t3 = std::chrono::steady_clock::now();
    std::memcpy(x, x2.data(), sizeof(double) * A.n);
    fillPerf(p, perf);
    seconds[0] = std::chrono::duration<double>(t1 - t0).count();
    seconds[1] = std::chrono::duration<double>(t2 - t1).count();
    seconds[2] = std::chrono::duration<double>(t3 - t2).count();
    return 0;
}

}  // extern "C"
