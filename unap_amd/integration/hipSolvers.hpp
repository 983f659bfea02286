// hipSolvers.hpp -- UNAP's solver, preconditioner and smoother classes backed
// by libunap_hip.so (include/unap_hip.h), for code built against the
// reference's own headers (src/matrix/unapMatrix.hpp, lduMatrix.hpp).
//
// A maintainer adds this file to the reference tree (for example as
// src/hip/hipSolvers.hpp) and links libunap_hip.so; callers keep their
// lduMatrix and their solve(x, A, b) calls and swap the class names:
//
//   UNAP::lduMatrix A(n, lowerAddr, upperAddr, lower, diag, upper);  // as before
//   UNAP::hip::deviceLdu dA(A);                    // device mirror, built once
//   UNAP::hip::precond P(dA, UNAP_HIP_PRECOND_DIC);  // for lduDICPrecond P(A)
//   UNAP::hip::krylov S(dA, UNAP_HIP_SOLVER_PCG, &P); // for PCG S(P)
//   S.SET_relTol(1e-8);
//   matrix::solverPerformance perf = S.solve(x, A, b);
//
// Vectors stay host scalarFields; each call copies them to the device and
// back.  Library errors throw std::runtime_error (the reference exits).
// Reference interfaces implemented: matrix::solver (unapMatrix.hpp:247-323),
// matrix::preconditioner (:326-345), matrix::smoother (:351-373).
#ifndef UNAP_HIP_SOLVERS_HPP
#define UNAP_HIP_SOLVERS_HPP

#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "lduMatrix.hpp"
#include "unapMatrix.hpp"
#include "unap_hip.h"

namespace UNAP {
namespace hip {

inline void check(int rc) {
    if (rc != UNAP_HIP_OK) throw std::runtime_error(std::string("unap_hip: ") + unap_hip_last_error());
}

// Device mirror of an lduMatrix (addressing and coefficients copied once).
// After the reference's setMatrixCoeffients on A, refresh() hands the new
// values to the device (AMG hierarchies then redo their Galerkin levels).
class deviceLdu {
    const lduMatrix& A_;
    unap_hip_ldu h_;

    deviceLdu(const deviceLdu&);
    deviceLdu& operator=(const deviceLdu&);

  public:
    explicit deviceLdu(const lduMatrix& A) : A_(A), h_(NULL) {
        const label nf = A.upperAddr().size();
        std::vector<int64_t> l(nf), u(nf);
        for (label f = 0; f < nf; ++f) {
            l[f] = A.lowerAddr()[f];
            u[f] = A.upperAddr()[f];
        }
        const double* lo = A.symm() ? NULL : A.lower().values();
        check(unap_hip_ldu_create(A.size(), nf, l.data(), u.data(), lo, A.diag().values(), A.upper().values(), &h_));
    }
    ~deviceLdu() { unap_hip_ldu_destroy(h_); }
    void refresh() {
        check(unap_hip_ldu_set_coeffs(h_, A_.diag().values(), A_.upper().values(),
                                      A_.symm() ? NULL : A_.lower().values()));
    }
    unap_hip_ldu get() const { return h_; }
    const lduMatrix& host() const { return A_; }
};

// matrix::preconditioner on the GPU: UNAP_HIP_PRECOND_DIAG / DIC / DILU
// (lduDiagPrecond, lduDICPrecond, lduDILUPrecond) or AMG (the MGPrecond
// composition), with AMG / precision parameters as unap_hip_precond_create.
class precond : public matrix::preconditioner {
    const deviceLdu& A_;
    unap_hip_precond h_;
    mutable scalarField rD_;

    precond(const precond&);
    precond& operator=(const precond&);

  public:
    precond(const deviceLdu& A, int type, const unap_hip_amg_params* amg = NULL, const double* weights = NULL)
        : A_(A), h_(NULL), rD_(A.host().size()) {
        check(unap_hip_precond_create(A.get(), type, amg, weights, &h_));
    }
    ~precond() { unap_hip_precond_destroy(h_); }
    virtual void precondition(scalarField& w, const scalarField& r) const {
        check(unap_hip_precond_apply(h_, r.values(), w.values(), UNAP_HIP_HOST));
    }
    virtual const scalarField& rD() const {
        check(unap_hip_precond_rd(h_, rD_.values()));
        return rD_;
    }
    unap_hip_precond get() const { return h_; }
};

inline matrix::solverPerformance toPerf(const unap_hip_perf& p) {
    matrix::solverPerformance out(p.initial_residual, p.final_residual, (label)p.n_iterations, p.converged != 0,
                                  p.singular != 0);
    out.previousResidual() = p.previous_residual;
    return out;
}

// matrix::solver on the GPU: UNAP_HIP_SOLVER_PCG / PBICGSTAB (PCG.cpp,
// PBiCGStab.cpp) or the flexible FCG / FBICGSTAB, around a precond (NULL:
// identity).  The limits are the SET_* values of this object at solve time.
class krylov : public matrix::solver {
    const deviceLdu& A_;
    unap_hip_solver h_;

    krylov(const krylov&);
    krylov& operator=(const krylov&);

  public:
    krylov(const deviceLdu& A, int type, const precond* P) : A_(A), h_(NULL) {
        check(unap_hip_solver_create(type, &h_));
        check(unap_hip_solver_set_precond(h_, P ? P->get() : NULL));
    }
    ~krylov() { unap_hip_solver_destroy(h_); }
    // A must be the lduMatrix the deviceLdu mirrors.
    virtual matrix::solverPerformance solve(scalarField& x, const matrix& A, const scalarField& b) const {
        if (&A != static_cast<const matrix*>(&A_.host()))
            throw std::runtime_error("unap_hip: solve() on a matrix other than the mirrored one");
        check(unap_hip_solver_set_limits(h_, maxIter(), minIter(), relTol(), tolerance()));
        unap_hip_perf p;
        check(unap_hip_solver_solve(h_, A_.get(), x.values(), b.values(), UNAP_HIP_HOST, &p));
        return toPerf(p);
    }
};

// MGSolver (MG.hpp, MGSolve.cpp) on the GPU: agglomeration, smoothers and
// sweep counts from the AMG parameters; face weights NULL: |upper| / ||upper||.
class MGSolver : public matrix::solver {
    const deviceLdu& A_;
    unap_hip_solver h_;

    MGSolver(const MGSolver&);
    MGSolver& operator=(const MGSolver&);

  public:
    MGSolver(const deviceLdu& A, const unap_hip_amg_params& amg, const double* weights = NULL) : A_(A), h_(NULL) {
        check(unap_hip_solver_create(UNAP_HIP_SOLVER_MG, &h_));
        check(unap_hip_solver_set_mg(h_, &amg, weights));
    }
    ~MGSolver() { unap_hip_solver_destroy(h_); }
    virtual matrix::solverPerformance solve(scalarField& x, const matrix& A, const scalarField& b) const {
        if (&A != static_cast<const matrix*>(&A_.host()))
            throw std::runtime_error("unap_hip: solve() on a matrix other than the mirrored one");
        check(unap_hip_solver_set_limits(h_, maxIter(), minIter(), relTol(), tolerance()));
        unap_hip_perf p;
        check(unap_hip_solver_solve(h_, A_.get(), x.values(), b.values(), UNAP_HIP_HOST, &p));
        return toPerf(p);
    }
};

// matrix::smoother on the GPU: UNAP_HIP_SMOOTHER_GS or _CHEBY
// (lduGaussSeidelSmoother, chebySmoother).  The smoother is bound to one
// deviceLdu, the mirror of the A passed to smooth().
class smoother : public matrix::smoother {
    const deviceLdu& A_;
    unap_hip_smoother h_;

    smoother(const smoother&);
    smoother& operator=(const smoother&);

  public:
    smoother(const deviceLdu& A, int type) : A_(A), h_(NULL) { check(unap_hip_smoother_create(type, &h_)); }
    ~smoother() { unap_hip_smoother_destroy(h_); }
    virtual void smooth(scalarField& x, const matrix& A, const scalarField& b, const label nSweeps) const {
        if (&A != static_cast<const matrix*>(&A_.host()))
            throw std::runtime_error("unap_hip: smooth() on a matrix other than the mirrored one");
        check(unap_hip_smoother_smooth(h_, A_.get(), x.values(), b.values(), nSweeps, UNAP_HIP_HOST));
    }
    virtual void init() const { check(unap_hip_smoother_init(h_)); }
};

}  // namespace hip
}  // namespace UNAP

#endif
