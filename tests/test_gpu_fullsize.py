"""The configurations at their full sizes, through properties that do not
depend on size (the CPU oracle would take minutes per solve here; the
bit-exact and history comparisons run at oracle-sized cases in
test_gpu_parity.py):

  * C2: 128^3 permuted log-uniform(1e-3..1) Poisson, PCG + FP64 AMG with GS
    smoothers, coarsest threshold 500, rel. tol 1e-8;
  * C3: 128^3 permuted convection-diffusion (Peclet up to 5), PBiCGStab +
    DILU, rel. tol 1e-8;
  * C4: 256^3 permuted log-uniform(1e-4..1) Poisson, the mixed FCG + FP32
    AMG path, rel. tol 1e-8.

Each must converge; its true residual ||b - A x|| (computed on the host from
the LDU arrays, independently of the GPU) must meet the tolerance the solver
reported; a second solve must reproduce the first bit for bit (the tree
reductions are deterministic); the residual history must decrease overall.
"""
import numpy as np
import pytest

from unap_amd import hip, meshgen

pytestmark = pytest.mark.gpu


def _true_rel_residual(s, x):
    y = s.diag * x
    y += np.bincount(s.l, weights=s.upper * x[s.u], minlength=s.n)
    y += np.bincount(s.u, weights=s.lower_or_upper() * x[s.l], minlength=s.n)
    r0 = np.linalg.norm(s.b)  # x0 = 0: the initial residual is b
    return np.linalg.norm(s.b - y) / r0


def _solve_twice(S, A, s):
    b = hip.DeviceArray.from_numpy(s.b)
    x = hip.DeviceArray(s.n)
    _, p1 = S.solve(A, b, x=x)
    x1 = x.numpy()
    x.zero()
    _, p2 = S.solve(A, b, x=x)
    assert np.array_equal(x.numpy(), x1), "second solve differs"
    assert p1["nIterations"] == p2["nIterations"]
    return x1, p1, S.history()


def test_c2_amg_pcg_gs_128():
    s = meshgen.config_system(2, 128)
    A = hip.LduMatrix.from_system(s)
    P = hip.Precond(A, "AMG", amg=hip.amg_params(smoother=1, n_cells_coarsest=500, n_vcycles=1))
    S = hip.Solver("PCG", P, rel_tol=1e-8, max_iter=1000)
    x, p, h = _solve_twice(S, A, s)
    assert p["converged"], p
    assert p["final"] <= 1e-8 * p["initial"]
    assert _true_rel_residual(s, x) <= 2e-8
    assert h[-1] < 1e-6 * h[0] and len(h) == p["nIterations"] + 1


def test_c3_pbicgstab_dilu_128():
    s = meshgen.config_system(3, 128)
    A = hip.LduMatrix.from_system(s)
    S = hip.Solver("PBiCGStab", hip.Precond(A, "DILU"), rel_tol=1e-8, max_iter=1000)
    x, p, h = _solve_twice(S, A, s)
    assert p["converged"], p
    assert _true_rel_residual(s, x) <= 2e-8


def test_c4_mixed_fcg_256():
    s = meshgen.config_system(4, 256)
    A = hip.LduMatrix.from_system(s)
    P = hip.Precond(A, "AMG", amg=hip.amg_params(precision=32, smoother=2, n_cells_coarsest=2000, n_vcycles=1,
                                                   n_post_sweeps=4, n_finest_sweeps=4, level_sweep_growth=0,
                                                   pair_steps=2))
    S = hip.Solver("FCG", P, rel_tol=1e-8, max_iter=1000)
    x, p, h = _solve_twice(S, A, s)
    assert p["converged"], p
    assert _true_rel_residual(s, x) <= 2e-8
    assert p["nIterations"] < 150


def test_c2_fp32_amg_gs_fcg_128():
    # the precision pattern on C2: FCG around the FP32 AMG with Gauss-Seidel
    # (fixed cycle, dense FP64 coarsest level)
    s = meshgen.config_system(2, 128)
    A = hip.LduMatrix.from_system(s)
    P = hip.Precond(A, "AMG", amg=hip.amg_params(precision=32, smoother=1, n_cells_coarsest=500, n_vcycles=1,
                                                   cycle=hip.CYCLE_FIXED))
    S = hip.Solver("FCG", P, rel_tol=1e-8, max_iter=1000)
    x, p, h = _solve_twice(S, A, s)
    assert p["converged"], p
    assert _true_rel_residual(s, x) <= 2e-8


def test_c3_bicgstab_dilu32_128():
    # the precision pattern on C3: PBiCGStab in FP64 around FP32 DILU sweeps
    s = meshgen.config_system(3, 128)
    A = hip.LduMatrix.from_system(s)
    S = hip.Solver("PBiCGStab", hip.Precond(A, "DILU", precision=32), rel_tol=1e-8, max_iter=2000)
    x, p, h = _solve_twice(S, A, s)
    assert p["converged"], p
    assert _true_rel_residual(s, x) <= 2e-8
