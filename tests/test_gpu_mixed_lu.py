"""FP32 DIC / DILU preconditioners inside FP64 Krylov solvers (FCG for DIC,
PBiCGStab or the flexible BiCGStab for DILU): the sweeps, their coefficients and rD in FP32, the
residual, dot products, norms and convergence test in FP64.  There is no
counterpart in the reference, so the checks are against the FP64 oracle:
the FP32 apply agrees with the FP64 apply to FP32 accuracy, each solve meets
its tolerance in the true residual, its iteration count stays near the FP64
one, and its error against an FP64 solution to 1e-13 is what the tolerance
allows (within 100x the FP64 solve's own error, or 1e-6: the error a given
residual allows scales with the conditioning)."""
import numpy as np
import pytest

from unap_amd import hip, meshgen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,c,N", [("DIC", 1, 12), ("DIC", 2, 12), ("DIC", 3, 12), ("DILU", 3, 12), ("DILU", 2, 10)])
def test_fp32_apply_matches_fp64(oracle, kind, c, N):
    s = meshgen.config_system(c, N)
    A = hip.LduMatrix.from_system(s)
    r = np.random.default_rng(7).uniform(-1, 1, s.n)
    w32 = hip.Precond(A, kind, precision=32).apply(r)
    w64, _ = oracle.matrix(s).precond(kind, r)
    assert np.linalg.norm(w32 - w64) <= 1e-5 * np.linalg.norm(w64)


def _check(oracle, s, solver, kind, rel_tol, x, p):
    M = oracle.matrix(s)
    assert p["converged"], p
    assert np.linalg.norm(s.b - M.spmv(x)) <= 2 * rel_tol * np.linalg.norm(s.b)
    fp64 = "PCG" if kind == "DIC" else "PBiCGStab"
    xo, po = M.solve(fp64, kind, s.b, rel_tol=rel_tol, max_iter=2000)
    assert abs(p["nIterations"] - po["nIterations"]) <= max(5, po["nIterations"] // 2), (p, po["nIterations"])
    x_ref, _ = M.solve(fp64, kind, s.b, rel_tol=1e-13, max_iter=5000)
    e32 = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    e64 = np.linalg.norm(xo - x_ref) / np.linalg.norm(x_ref)
    assert e32 <= max(100 * e64, 1e-6), (e32, e64)


@pytest.mark.parametrize("c,N,tol", [(1, 32, 1e-6), (1, 16, 1e-8), (2, 16, 1e-8)])
def test_fcg_dic32(oracle, c, N, tol):
    s = meshgen.config_system(c, N)
    A = hip.LduMatrix.from_system(s)
    S = hip.Solver("FCG", hip.Precond(A, "DIC", precision=32), rel_tol=tol, max_iter=2000)
    x, p = S.solve(A, s.b)
    _check(oracle, s, "FCG", "DIC", tol, x, p)


@pytest.mark.parametrize("solver", ["PBiCGStab", "FBiCGStab"])
@pytest.mark.parametrize("c,N", [(3, 16), (3, 24), (2, 16)])
def test_bicgstab_dilu32(oracle, solver, c, N):
    s = meshgen.config_system(c, N)
    A = hip.LduMatrix.from_system(s)
    S = hip.Solver(solver, hip.Precond(A, "DILU", precision=32), rel_tol=1e-8, max_iter=2000)
    x, p = S.solve(A, s.b)
    _check(oracle, s, solver, "DILU", 1e-8, x, p)


def test_fp32_lu_needs_flexible_solver():
    s = meshgen.config_system(1, 8)
    A = hip.LduMatrix.from_system(s)
    S = hip.Solver("PCG", hip.Precond(A, "DIC", precision=32))
    with pytest.raises(hip.UnapHipError):
        S.solve(A, s.b)
