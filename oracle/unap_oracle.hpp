// unap_oracle.hpp -- CPU restatement of UNAP's solve path.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU library is
// compared against; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  Nothing in the product (unap_amd/) links,
// imports or calls it.
//
// Every routine restates the serial x86 branch of the reference (file:line
// cited in unap_oracle.cpp) with the same floating-point operation order, and
// is built with -O3 -ffp-contract=off so no fused multiply-add changes a bit.
// Parity of each routine is pinned by tests/test_oracle_*.py against the
// compiled reference (oracle/_ref), the committed golden vectors in
// tests/golden/, and closed forms / invariants listed there.
#pragma once

#include <cstdint>
#include <vector>

namespace unap_oracle {

typedef int64_t label;

// matrix::solverPerformance (src/matrix/unapMatrix.hpp:43-174)
struct Perf {
    double initial = 0, final_ = 0, previous = 0;
    label nIter = 0;
    bool converged = false, singular = false;
};

bool checkConvergence(Perf& p, double tol, double relTol, label iter, label minIter);
bool checkSingularity(Perf& p, double residual);

// lduMatrix (src/matrix/lduMatrix/lduMatrix.hpp:17-349)
struct Ldu {
    label n = 0;
    std::vector<label> l, u;        // lowerAddr (owner), upperAddr (neighbour)
    std::vector<double> lo, up, d;  // lower (empty when symmetric), upper, diag
    bool symm = true;

    label nFaces() const { return (label)u.size(); }
    const double* upper() const { return up.data(); }
    const double* lower() const { return symm ? up.data() : lo.data(); }

    void spmv(double* y, const double* x) const;
    const std::vector<label>& ownerStart() const;
    const std::vector<label>& losort() const;
    const std::vector<label>& losortStart() const;

  private:
    mutable std::vector<label> ownerStart_, losort_, losortStart_;
};

double dot(const double* a, const double* b, label n);
double sumSqr(const double* a, label n);
double normFactor(const double* a, label n);

// matrix::preconditioner and the ldu preconditioners
struct Precond {
    virtual ~Precond() {}
    virtual void apply(double* w, const double* r, label n) const;  // identity
    virtual const double* rD() const { return nullptr; }
};

struct DiagPrecond : Precond {
    std::vector<double> rd;
    explicit DiagPrecond(const Ldu& A);
    void apply(double* w, const double* r, label n) const override;
    const double* rD() const override { return rd.data(); }
};

struct DICPrecond : Precond {
    const Ldu* A;
    std::vector<double> rd;
    explicit DICPrecond(const Ldu& A);
    void apply(double* w, const double* r, label n) const override;
    const double* rD() const override { return rd.data(); }
};

struct DILUPrecond : Precond {
    const Ldu* A;
    std::vector<double> rd;
    explicit DILUPrecond(const Ldu& A);
    void apply(double* w, const double* r, label n) const override;
    const double* rD() const override { return rd.data(); }
};

// One forward Gauss-Seidel sweep from zero as a preconditioner,
// w = (D + L)^-1 r: a linear but non-symmetric preconditioner built from the
// reference's smoother (lduGaussSeidelSmoother.cpp:3-80).  Test-only: it
// exercises the flexible-CG direction update, which a symmetric
// preconditioner cannot tell apart from PCG's.
struct GSPrecond : Precond {
    const Ldu* A;
    explicit GSPrecond(const Ldu& a) : A(&a) {}
    void apply(double* w, const double* r, label n) const override;
};

// smoothers
struct Smoother {
    virtual ~Smoother() {}
    virtual void smooth(double* x, const Ldu& A, const double* b, label nSweeps) const = 0;
    virtual void init() const {}
};

struct GSSmoother : Smoother {
    void smooth(double* x, const Ldu& A, const double* b, label nSweeps) const override;
};

// eigenDiagPCG: returns the estimate and (optionally) the recorded alphas/betas.
double eigenDiagPCG(const Ldu& A, double* x, const double* b, const double* rD, label nPCGs,
                    std::vector<double>* alphas = nullptr, std::vector<double>* betas = nullptr);
double maxEigFromLanczos(const std::vector<double>& alphas, const std::vector<double>& betas);

struct ChebySmoother : Smoother {
    label nDiagPCGs = 10;
    double eigRatio = 1.5, boost = 1.1;
    mutable bool firstCall = true;
    mutable double maxEig = 0;
    void smooth(double* x, const Ldu& A, const double* b, label nSweeps) const override;
    void init() const override { firstCall = true; }
};

// lduAgglomeration
struct Agglomeration {
    const Ldu* fine;
    label maxLevels = 50, nCoarsest = 50;
    std::vector<label> nCells;
    std::vector<Ldu> coarse;  // topology, then Galerkin coefficients
    std::vector<std::vector<label>> restrictAddr, faceRestrictAddr;

    explicit Agglomeration(const Ldu& A) : fine(&A) {}
    label size() const { return (label)coarse.size(); }
    const Ldu& matrixLevel(label i) const { return i == 0 ? *fine : coarse[i - 1]; }
    void agglomerate(const double* weights);
    void agglomerateMatrix(label fineLevel);
    void agglomerateAll() {
        for (label i = 0; i < size(); ++i) agglomerateMatrix(i);
    }
    void restrictField(double* cf, const double* ff, label fineLevel) const;
    void prolongField(double* ff, const double* cf, label coarseLevel) const;
    void restrictFaceField(double* cf, const double* ff, label fineLevel) const;

  private:
    std::vector<label> pairLevel(label& nCoarse, const Ldu& fineA, const double* w) const;
    void buildCoarseAddressing(label fineLevel);
};

// solvers
struct Solver {
    label maxIter = 1000, minIter = 0;
    double relTol = 1e-6, tol = 0;
    std::vector<double>* history = nullptr;  // initial residual, then one per iteration
    virtual ~Solver() {}
    virtual Perf solve(double* x, const Ldu& A, const double* b) const = 0;
};

struct PCG : Solver {
    const Precond* M;
    explicit PCG(const Precond& m) : M(&m) {}
    Perf solve(double* x, const Ldu& A, const double* b) const override;
};

struct PBiCGStab : Solver {
    const Precond* M;
    // 0: the reference's PBiCGStab.  > 0: the flexible variant (no reference
    // counterpart), which restarts the shadow residual, r0 = r, once
    // |r0.r| < shadowRestart |r0| |r| (then p = r as in a first iteration).
    double shadowRestart = 0;
    explicit PBiCGStab(const Precond& m) : M(&m) {}
    Perf solve(double* x, const Ldu& A, const double* b) const override;
};

struct FBiCGStab : PBiCGStab {
    explicit FBiCGStab(const Precond& m) : PBiCGStab(m) { shadowRestart = 1.0e-6; }
};

// Flexible CG (Notay 2000 FCG(1): each direction A-orthogonalised against the
// previous one): the outer loop of the mixed-precision path; no counterpart in
// the reference.
struct FCG : Solver {
    const Precond* M;
    explicit FCG(const Precond& m) : M(&m) {}
    Perf solve(double* x, const Ldu& A, const double* b) const override;
};

struct MGSolver : Solver {
    const Ldu* finest;
    Agglomeration* ag;
    std::vector<const Smoother*> sm;
    label nPre = 0, nPost = 2, nFinest = 2;
    bool scale;
    MGSolver(const Ldu& A, Agglomeration& a, const std::vector<const Smoother*>& s)
        : finest(&A), ag(&a), sm(s), scale(A.symm) {}
    Perf solve(double* x, const Ldu& A, const double* b) const override;
    double scalingFactor(double* Acf, double* field, const double* source, const Ldu& A) const;
    void vcycle(double* psi, const double* source, double* Apsi, double* finestCorr,
                double* finestRes, std::vector<std::vector<double>>& corr,
                std::vector<std::vector<double>>& src) const;
    void solveCoarsest(double* corr, const double* src) const;
};

// The agglomeration-AMG preconditioner: the reference's MGPrecond is a stub,
// this is the composition it sketches: w = 0, then MGSolver::solve(w, A, r)
// running exactly nVcycles V-cycles.
struct MGPrecond : Precond {
    const Ldu* A;
    Agglomeration ag;
    std::vector<Smoother*> owned;
    MGSolver* mg;
    label nVcycles;
    // smootherType 1 = Gauss-Seidel, 2 = Chebyshev
    MGPrecond(const Ldu& A, int smootherType, label nCoarsest, label nVcycles, label nPre,
              label nPost, label nFinest, const double* weights = nullptr);
    ~MGPrecond() override;
    void apply(double* w, const double* r, label n) const override;
    void init() const {
        for (auto* s : owned) s->init();
    }
};

}  // namespace unap_oracle
