// Drives the C++ API (include/unap_hip.hpp) the way the reference's examples
// drive UNAP's (test/ex3.cpp, doc/latex/CG-example.cpp): build an LDU matrix,
// a preconditioner and a solver, then solve(x, A, b).  Prints one line per
// check and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "unap_hip.hpp"
#include "unap_meshgen.h"

using namespace unap::hip;

static int fail(const char* what) {
    std::printf("FAIL %s\n", what);
    return 1;
}

int main() {
    // test/ex3.cpp: A = [[3,2],[2,6]], b = [2,-8] -> x = [2,-2] with PCG + DIC
    {
        std::vector<long> lAddr{0}, uAddr{1};
        std::vector<double> lower{2.0}, upper{2.0}, diag{3.0, 6.0}, b{2.0, -8.0}, x(2, 0.0);
        lduMatrix A(2, lAddr, uAddr, lower, diag, upper);
        DICPrecond P(A);
        PCG S(P);
        solverPerformance perf = S.solve(x, A, b);
        std::printf("ex3: %ld iterations, x = [%.15g, %.15g]\n", (long)perf.nIterations(), x[0], x[1]);
        if (std::fabs(x[0] - 2.0) > 1e-12 || std::fabs(x[1] + 2.0) > 1e-12) return fail("ex3");
    }
    // a permuted 16^3 log-uniform Poisson system: PCG+DIC, PBiCGStab+DILU,
    // AMG-PCG, the MG solver, and FCG / PBiCGStab around FP32 DIC / DILU all
    // reach relTol 1e-8
    const int64_t N = 16, n = N * N * N, nf = unap_meshgen_nfaces(N, N, N);
    std::vector<int64_t> l(nf), u(nf);
    std::vector<double> up(nf), d(n), b(n);
    unap_meshgen_box(N, N, N, UNAP_MESHGEN_POISSON_LOGUNIF, -3.0, 0.0, 1e-6, 1002, l.data(), u.data(), nullptr,
                     d.data(), up.data(), b.data(), nullptr);
    lduMatrix A(n, l, u, up, d, up);
    if (!A.symm()) return fail("symmetric flag");
    std::vector<double> y(n), ones(n, 1.0);
    A.spMV(y, ones);  // A 1 = row sums = the 1e-6 shift of every diagonal
    for (int64_t i = 0; i < n; ++i)
        if (std::fabs(y[i] - 1e-6) > 1e-12) return fail("A*1 row sums");
    DICPrecond dic(A);
    DILUPrecond dilu(A);
    unap_hip_amg_params prm = defaultAMG();
    prm.smoother = UNAP_HIP_SMOOTHER_CHEBY;
    prm.n_vcycles = 1;
    MGPrecond amg(A, prm);
    DICPrecond dic32(A, 32);
    DILUPrecond dilu32(A, 32);
    PCG s1(dic);
    PBiCGStab s2(dilu);
    PCG s3(amg);
    MGSolver s4(prm);
    FCG s5(dic32);
    PBiCGStab s6(dilu32);
    solver* all[] = {&s1, &s2, &s3, &s4, &s5, &s6};
    const char* names[] = {"PCG+DIC", "PBiCGStab+DILU", "AMG-PCG", "MGSolver", "FCG+DIC(FP32)",
                           "PBiCGStab+DILU(FP32)"};
    for (int k = 0; k < 6; ++k) {
        all[k]->SET_relTol(1e-8);
        all[k]->SET_maxIter(500);
        std::vector<double> x(n, 0.0);
        solverPerformance p = all[k]->solve(x, A, b);
        A.spMV(y, x);
        double rr = 0, bb = 0;
        for (int64_t i = 0; i < n; ++i) {
            rr += (b[i] - y[i]) * (b[i] - y[i]);
            bb += b[i] * b[i];
        }
        std::printf("%s: %ld iterations, converged %d, true rel. residual %.3e, history %zu\n", names[k],
                    (long)p.nIterations(), (int)p.converged(), std::sqrt(rr / bb), all[k]->residualHistory().size());
        if (!p.converged() || std::sqrt(rr / bb) > 1.5e-8) return fail(names[k]);
    }
    // errors come back as exceptions, not exits
    try {
        std::vector<long> badL{0}, badU{5};
        std::vector<double> one{1.0}, dd{1.0, 1.0};
        lduMatrix bad(2, badL, badU, one, dd, one);
        return fail("bad address accepted");
    } catch (const error& e) {
        std::printf("bad input rejected: code %d (%s)\n", e.code, e.what());
    }
    std::printf("OK\n");
    return 0;
}
