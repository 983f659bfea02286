// unap_hip.hpp -- C++ classes over the C-ABI (unap_hip.h) with the shape of
// the reference's plug-in API (src/matrix/unapMatrix.hpp:9-375):
//
//   UNAP reference                         this header (namespace unap::hip)
//   lduMatrix(n, lAddr, uAddr, lo, d, up)  lduMatrix(n, lAddr, uAddr, lo, d, up)
//   A.spMV(Apsi, psi)                      A.spMV(Apsi, psi)
//   lduDiagPrecond / lduDICPrecond /       diagPrecond / DICPrecond /
//   lduDILUPrecond (A)                     DILUPrecond (A), MGPrecond(A, params)
//   P.precondition(w, r), P.rD()           P.precondition(w, r), P.rD(out)
//   lduGaussSeidelSmoother, chebySmoother  GaussSeidelSmoother, chebySmoother
//   S.smooth(x, A, b, nSweeps), S.init()   S.smooth(x, A, b, nSweeps), S.init()
//   lduAgglomeration(A).agglomerate(w)     agglomeration(A, w, maxLevels, nCoarsest)
//   PCG(P) / PBiCGStab(P) / MGSolver       PCG(P) / PBiCGStab(P) / MGSolver(params) / FCG(P)
//   S.SET_maxIter / minIter / relTol / tolerance
//   solverPerformance S.solve(x, A, b)     solverPerformance S.solve(x, A, b)
//
// Fields are any contiguous container with size() and data() / values()
// (std::vector<double>, or UNAP's scalarField / labelField, whose values()
// gives the array).  Errors throw unap::hip::error with the library's status
// code and message instead of exiting.  Objects own their handles (not
// copyable, movable).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "unap_hip.h"

namespace unap {
namespace hip {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != UNAP_HIP_OK) throw error(rc, unap_hip_last_error());
}

namespace detail {
template <class F>
auto ptr(F& f, int) -> decltype(f.data()) {
    return f.data();
}
template <class F>
auto ptr(F& f, long) -> decltype(f.values()) {
    return f.values();
}
template <class F>
auto p(F& f) -> decltype(ptr(f, 0)) {
    return ptr(f, 0);
}
template <class F>
int64_t len(const F& f) {
    return (int64_t)f.size();
}
// labels of any integer width -> int64
template <class F>
std::vector<int64_t> labels(const F& f) {
    std::vector<int64_t> out((size_t)f.size());
    auto* s = p(const_cast<F&>(f));
    for (size_t i = 0; i < out.size(); ++i) out[i] = (int64_t)s[i];
    return out;
}
}  // namespace detail

// solverPerformance (unapMatrix.hpp:43-174)
struct solverPerformance {
    unap_hip_perf v{};
    double initialResidual() const { return v.initial_residual; }
    double finalResidual() const { return v.final_residual; }
    double previousResidual() const { return v.previous_residual; }
    int64_t nIterations() const { return v.n_iterations; }
    bool converged() const { return v.converged != 0; }
    bool singular() const { return v.singular != 0; }
};

template <class H, int (*Destroy)(H)>
class handle {
  protected:
    H h_ = nullptr;

  public:
    handle() = default;
    handle(const handle&) = delete;
    handle& operator=(const handle&) = delete;
    handle(handle&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    handle& operator=(handle&& o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    ~handle() {
        if (h_) Destroy(h_);
    }
    H get() const { return h_; }
};

class lduMatrix : public handle<unap_hip_ldu, unap_hip_ldu_destroy> {
    int64_t n_ = 0;

  public:
    // lduMatrix.cpp:28-68; pass the same object as lower and upper for a
    // symmetric matrix (one coefficient array), as the reference does.
    template <class LF, class SF>
    lduMatrix(int64_t nCells, const LF& lowerAddr, const LF& upperAddr, const SF& lower, const SF& diag,
              const SF& upper)
        : n_(nCells) {
        const std::vector<int64_t> la = detail::labels(lowerAddr), ua = detail::labels(upperAddr);
        const double* up = detail::p(const_cast<SF&>(upper));
        const double* lo = (&lower == &upper) ? nullptr : detail::p(const_cast<SF&>(lower));
        check(unap_hip_ldu_create(nCells, (int64_t)la.size(), la.data(), ua.data(), lo,
                                  detail::p(const_cast<SF&>(diag)), up, &h_));
    }
    int64_t size() const { return n_; }
    bool symm() const {
        int s = 0;
        check(unap_hip_ldu_info(h_, nullptr, nullptr, &s, nullptr, nullptr));
        return s != 0;
    }
    template <class SF>
    void spMV(SF& Apsi, const SF& psi) const {
        check(unap_hip_ldu_spmv(h_, detail::p(const_cast<SF&>(psi)), detail::p(Apsi), UNAP_HIP_HOST));
    }
    template <class SF>
    void setMatrixCoeffients(const SF& diag, const SF& upper, const SF& lower) {
        const double* lo = (&lower == &upper) ? nullptr : detail::p(const_cast<SF&>(lower));
        check(unap_hip_ldu_set_coeffs(h_, detail::p(const_cast<SF&>(diag)), detail::p(const_cast<SF&>(upper)), lo));
    }
};

class preconditioner : public handle<unap_hip_precond, unap_hip_precond_destroy> {
  protected:
    int64_t n_ = 0;
    preconditioner(const lduMatrix& A, int type, const unap_hip_amg_params* amg = nullptr,
                   const double* weights = nullptr)
        : n_(A.size()) {
        check(unap_hip_precond_create(A.get(), type, amg, weights, &h_));
    }

    // parameters that carry only a precision (read by DIC / DILU)
    static const unap_hip_amg_params* withPrecision(int precision) {
        static thread_local unap_hip_amg_params p;
        check(unap_hip_amg_params_default(&p));
        p.precision = precision;
        return &p;
    }

  public:
    template <class SF>
    void precondition(SF& w, const SF& r) const {
        check(unap_hip_precond_apply(h_, detail::p(const_cast<SF&>(r)), detail::p(w), UNAP_HIP_HOST));
    }
    template <class SF>
    void rD(SF& out) const {
        check(unap_hip_precond_rd(h_, detail::p(out)));
    }
};

inline unap_hip_amg_params defaultAMG() {
    unap_hip_amg_params p;
    check(unap_hip_amg_params_default(&p));
    return p;
}

struct identityPrecond : preconditioner {
    explicit identityPrecond(const lduMatrix& A) : preconditioner(A, UNAP_HIP_PRECOND_NONE) {}
};
struct diagPrecond : preconditioner {
    explicit diagPrecond(const lduMatrix& A) : preconditioner(A, UNAP_HIP_PRECOND_DIAG) {}
};
// precision 32: FP32 sweeps, for the FCG and PBiCGStab solvers (no
// counterpart in the reference); 64: the reference's arithmetic.
struct DICPrecond : preconditioner {
    explicit DICPrecond(const lduMatrix& A, int precision = 64)
        : preconditioner(A, UNAP_HIP_PRECOND_DIC, withPrecision(precision)) {}
};
struct DILUPrecond : preconditioner {
    explicit DILUPrecond(const lduMatrix& A, int precision = 64)
        : preconditioner(A, UNAP_HIP_PRECOND_DILU, withPrecision(precision)) {}
};

// The agglomeration-AMG preconditioner (MGPreconditioner.hpp:11-49 defines it
// as a stub; here precondition = nVcycles V-cycles of MGSolver from w = 0).
struct MGPrecond : preconditioner {
    explicit MGPrecond(const lduMatrix& A, const unap_hip_amg_params& p = defaultAMG())
        : preconditioner(A, UNAP_HIP_PRECOND_AMG, &p) {}
};

class smoother : public handle<unap_hip_smoother, unap_hip_smoother_destroy> {
  protected:
    explicit smoother(int type) { check(unap_hip_smoother_create(type, &h_)); }

  public:
    template <class SF>
    void smooth(SF& x, const lduMatrix& A, const SF& b, int64_t nSweeps) const {
        check(unap_hip_smoother_smooth(h_, A.get(), detail::p(x), detail::p(const_cast<SF&>(b)), nSweeps,
                                       UNAP_HIP_HOST));
    }
    void init() const { check(unap_hip_smoother_init(h_)); }
};
struct GaussSeidelSmoother : smoother {
    GaussSeidelSmoother() : smoother(UNAP_HIP_SMOOTHER_GS) {}
};
struct chebySmoother : smoother {
    chebySmoother() : smoother(UNAP_HIP_SMOOTHER_CHEBY) {}
};

class agglomeration : public handle<unap_hip_aggl, unap_hip_aggl_destroy> {
  public:
    template <class SF>
    agglomeration(const lduMatrix& A, const SF& weights, int64_t maxLevels = 50, int64_t nCellsInCoarsestLevel = 50) {
        check(unap_hip_aggl_create(A.get(), detail::p(const_cast<SF&>(weights)), maxLevels, nCellsInCoarsestLevel,
                                   &h_));
    }
    int64_t size() const {
        int64_t n = 0;
        check(unap_hip_aggl_nlevels(h_, &n));
        return n;
    }
    std::vector<int64_t> restrictAddressing(int64_t level, int64_t fineCells) const {
        std::vector<int64_t> out((size_t)fineCells);
        check(unap_hip_aggl_restrict_addressing(h_, level, out.data()));
        return out;
    }
};

class solver : public handle<unap_hip_solver, unap_hip_solver_destroy> {
  protected:
    explicit solver(int type, const preconditioner* P = nullptr) {
        check(unap_hip_solver_create(type, &h_));
        if (P) check(unap_hip_solver_set_precond(h_, P->get()));
        limits();
    }
    int64_t maxIter_ = 1000, minIter_ = 0;
    double relTol_ = 1e-6, tol_ = 0.0;
    void limits() { check(unap_hip_solver_set_limits(h_, maxIter_, minIter_, relTol_, tol_)); }

  public:
    // unapMatrixSolver.cpp:18-41
    void SET_maxIter(int64_t v) { maxIter_ = v, limits(); }
    void SET_minIter(int64_t v) { minIter_ = v, limits(); }
    void SET_relTol(double v) { relTol_ = v, limits(); }
    void SET_tolerance(double v) { tol_ = v, limits(); }
    template <class SF>
    solverPerformance solve(SF& x, const lduMatrix& A, const SF& b) const {
        solverPerformance p;
        check(unap_hip_solver_solve(h_, A.get(), detail::p(x), detail::p(const_cast<SF&>(b)), UNAP_HIP_HOST, &p.v));
        return p;
    }
    std::vector<double> residualHistory() const {
        int64_t n = 0;
        check(unap_hip_solver_history(h_, nullptr, 0, &n));
        std::vector<double> h((size_t)n);
        check(unap_hip_solver_history(h_, h.data(), n, &n));
        return h;
    }
};

struct PCG : solver {
    PCG() : solver(UNAP_HIP_SOLVER_PCG) {}
    explicit PCG(const preconditioner& P) : solver(UNAP_HIP_SOLVER_PCG, &P) {}
};
struct PBiCGStab : solver {
    PBiCGStab() : solver(UNAP_HIP_SOLVER_PBICGSTAB) {}
    explicit PBiCGStab(const preconditioner& P) : solver(UNAP_HIP_SOLVER_PBICGSTAB, &P) {}
};
struct FCG : solver {
    explicit FCG(const preconditioner& P) : solver(UNAP_HIP_SOLVER_FCG, &P) {}
};
struct MGSolver : solver {
    explicit MGSolver(const unap_hip_amg_params& p = defaultAMG()) : solver(UNAP_HIP_SOLVER_MG) {
        check(unap_hip_solver_set_mg(h_, &p, nullptr));
    }
};

}  // namespace hip
}  // namespace unap
