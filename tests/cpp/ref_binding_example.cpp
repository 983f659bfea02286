// ref_binding_example.cpp -- the reference's own classes beside the HIP ones of
// unap_amd/integration/hipSolvers.hpp on the same lduMatrix objects.
//
// Built against the reference's headers and sources (oracle/Makefile.ref,
// test infrastructure), run on a GPU by tests/test_gpu_ref_binding.py.  With
// reference-order reductions every HIP solve must reproduce the x86 solve bit
// for bit: iterations, final residual and solution.  Cases: PCG + DIC and
// PBiCGStab + DILU (C2 / C3 coefficients), MGSolver with Gauss-Seidel
// smoothers, and the same solver objects again after setMatrixCoeffients.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "MG.hpp"
#include "PBiCGStab.hpp"
#include "PCG.hpp"
#include "hipSolvers.hpp"
#include "lduAgglomeration.hpp"
#include "lduDICPrecond.hpp"
#include "lduDILUPrecond.hpp"
#include "lduGaussSeidelSmoother.hpp"
#include "unapMPI.hpp"
#include "unap_meshgen.h"

using namespace UNAP;

struct System {
    int64_t n = 0, nf = 0;
    std::vector<int64_t> l, u;
    std::vector<double> lo, d, up, b;
};

static System make(int kind, int N) {
    System s;
    s.n = unap_meshgen_ncells(N, N, N);
    s.nf = unap_meshgen_nfaces(N, N, N);
    s.l.resize(s.nf);
    s.u.resize(s.nf);
    s.up.resize(s.nf);
    s.d.resize(s.n);
    s.b.resize(s.n);
    if (kind == 3) s.lo.resize(s.nf);
    // C2: log-uniform 1e-3..1 Poisson; C3: convection-diffusion, Peclet 5
    const double p0 = kind == 2 ? -3.0 : 5.0, p1 = 0.0, shift = kind == 2 ? 1e-6 : 1e-3;
    unap_meshgen_box(N, N, N, kind, p0, p1, shift, 1000 + kind, s.l.data(), s.u.data(),
                     kind == 3 ? s.lo.data() : NULL, s.d.data(), s.up.data(), s.b.data(), NULL);
    return s;
}

static lduMatrix* toLdu(const System& s) {
    labelField la(s.nf), ua(s.nf);
    for (int64_t f = 0; f < s.nf; ++f) {
        la[f] = s.l[f];
        ua[f] = s.u[f];
    }
    scalarField dg(s.d.data(), s.n), up(s.up.data(), s.nf);
    if (s.lo.empty()) return new lduMatrix(s.n, la, ua, up, dg, up);
    scalarField lo(s.lo.data(), s.nf);
    return new lduMatrix(s.n, la, ua, lo, dg, up);
}

static int failures = 0;

static void compare(const char* what, const matrix::solver& ref, const matrix::solver& gpu, const lduMatrix& A,
                    const std::vector<double>& b) {
    scalarField bb(b.data(), (label)b.size()), x1((label)b.size(), 0.0), x2((label)b.size(), 0.0);
    matrix::solverPerformance p1 = ref.solve(x1, A, bb);
    matrix::solverPerformance p2 = gpu.solve(x2, A, bb);
    const bool same = p1.nIterations() == p2.nIterations() && p1.finalResidual() == p2.finalResidual() &&
                      std::memcmp(x1.values(), x2.values(), sizeof(double) * b.size()) == 0;
    std::printf("%-36s x86 %4ld it %.6e | HIP %4ld it %.6e | %s\n", what, (long)p1.nIterations(), p1.finalResidual(),
                (long)p2.nIterations(), p2.finalResidual(), same ? "identical" : "DIFFER");
    if (!same || !p1.converged()) ++failures;
}

int main() {
    unapMPI::initMPI();
    hip::check(unap_hip_set_reduction_order(UNAP_HIP_REDUCE_REFERENCE));
    {
        System s = make(2, 20);
        lduMatrix* A = toLdu(s);
        hip::deviceLdu dA(*A);
        lduDICPrecond rP(*A);
        PCG rS(rP);
        hip::precond hP(dA, UNAP_HIP_PRECOND_DIC);
        hip::krylov hS(dA, UNAP_HIP_SOLVER_PCG, &hP);
        rS.SET_relTol(1e-8);
        hS.SET_relTol(1e-8);
        compare("PCG + DIC, C2 20^3", rS, hS, *A, s.b);
        // MGSolver with Gauss-Seidel smoothers, weights |upper| / ||upper||
        scalarField w(s.nf);
        const double nrm = A->upper().SumSqrt();
        for (int64_t f = 0; f < s.nf; ++f) w[f] = std::fabs(A->upper()[f]) / nrm;
        lduAgglomeration ag(*A);
        ag.SET_nCellsInCoarsestLevel(20);
        ag.agglomerate(w);
        PtrList<matrix::smoother> sm(ag.size());
        for (label i = 0; i < ag.size(); ++i) sm.setLevel(i, *new lduGaussSeidelSmoother);
        MGSolver rM(*A, ag, sm);
        rM.SET_relTol(1e-8);
        rM.SET_maxIter(100);
        unap_hip_amg_params prm;
        hip::check(unap_hip_amg_params_default(&prm));
        prm.n_cells_coarsest = 20;
        hip::MGSolver hM(dA, prm);
        hM.SET_relTol(1e-8);
        hM.SET_maxIter(100);
        compare("MGSolver (Gauss-Seidel), C2 20^3", rM, hM, *A, s.b);
        // new coefficients, same solver objects (lduMatrix.cpp:508-529)
        std::vector<double> up2(s.up);
        for (int64_t f = 0; f < s.nf; ++f) up2[f] *= 0.5 + 0.5 * std::fabs(std::sin(0.37 * f));
        scalarField dg(s.d.data(), s.n), up(up2.data(), s.nf);
        A->setMatrixCoeffients(dg, up, up);
        dA.refresh();
        compare("PCG + DIC after setMatrixCoeffients", rS, hS, *A, s.b);
        compare("MGSolver after setMatrixCoeffients", rM, hM, *A, s.b);
        delete A;
    }
    {
        System s = make(3, 16);
        lduMatrix* A = toLdu(s);
        hip::deviceLdu dA(*A);
        lduDILUPrecond rP(*A);
        PBiCGStab rS(rP);
        hip::precond hP(dA, UNAP_HIP_PRECOND_DILU);
        hip::krylov hS(dA, UNAP_HIP_SOLVER_PBICGSTAB, &hP);
        rS.SET_relTol(1e-8);
        hS.SET_relTol(1e-8);
        compare("PBiCGStab + DILU, C3 16^3", rS, hS, *A, s.b);
        delete A;
    }
    std::printf("%s\n", failures ? "FAILED" : "ALL IDENTICAL");
    return failures ? 1 : 0;
}
