"""The FP64 drop-in against the oracle at the configurations' own sizes.

Tree reductions (the default) at 128^3, C2 and C3:
  * the initial residual agrees to 1e-11 relative: the Amul is bit-exact and
    only the order of the 2.1M-term sum differs (its rounding is bounded by
    n * eps ~ 4.7e-10, and is about sqrt(n) * eps in practice);
  * both converge, with iteration counts within 10 % (at least 1): a
    reordered sum perturbs every later iterate, and on these ill-conditioned
    systems the reference's own solver moves that much under a reordering
    (tests/test_oracle_pinning.py::test_history_sensitivity);
  * the residual history agrees to 1e-6 over the first quarter of the run
    (CG) or its first 20 iterations (BiCGStab): the reference's own solvers
    leave that band later than that under a reordering of their sums, and
    BiCGStab after a roughly fixed number of iterations, 26 at 32^3 and 36
    at 64^3, having stayed within 1e-8 for the first 20
    (tests/test_oracle_pinning.py::test_bicgstab_history_sensitivity);
  * the GPU solution is as accurate as the oracle's: its error against a
    tighter solution x* (relative residual 1e-11, checked on the host from
    the LDU arrays, so independent of who computed it) is at most 10x the
    oracle's error at the same tolerance.
The iteration delta and both errors are printed.

Reference-order reductions at 64^3 (the largest size whose serial sums keep
the suite within minutes): iterations, the whole residual history and the
solution bit for bit, for C1-type PCG+DIC, C2 AMG-PCG with Gauss-Seidel, C3
PBiCGStab+DILU and the C4 FP64 drop-in (PCG + AMG with Chebyshev smoothers).

At 256^3, the C4 FP64 drop-in (the denominator of the mixed-precision
speed-up) through size-independent properties: convergence, the true
residual on the host, and run-to-run bitwise determinism.
"""
import numpy as np
import pytest

from unap_amd import hip, meshgen

pytestmark = pytest.mark.gpu


def _host_amul(s, x):
    y = s.diag * x
    y += np.bincount(s.l, weights=s.upper * x[s.u], minlength=s.n)
    y += np.bincount(s.u, weights=s.lower_or_upper() * x[s.l], minlength=s.n)
    return y


def _rel_residual(s, x):
    return np.linalg.norm(s.b - _host_amul(s, x)) / np.linalg.norm(s.b)  # x0 = 0


def _gpu_solve(A, P, kind, b, tol, max_iter=2000):
    S = hip.Solver(kind, P, rel_tol=tol, max_iter=max_iter)
    return S.solve(A, b)


TREE_CASES = [
    ("C2 PCG+AMG(GS) 128^3", 2, "PCG", "AMG", [1, 500, 1, 0, 2, 2]),
    ("C3 PBiCGStab+DILU 128^3", 3, "PBiCGStab", "DILU", None),
]


@pytest.mark.parametrize("name,c,solver,pre,mg", TREE_CASES, ids=[t[0] for t in TREE_CASES])
def test_config_128_tree_vs_oracle(oracle, name, c, solver, pre, mg):
    s = meshgen.config_system(c, 128)
    A = hip.LduMatrix.from_system(s)
    amg = None if mg is None else hip.amg_params(smoother=mg[0], n_cells_coarsest=mg[1], n_vcycles=mg[2])
    P = hip.Precond(A, pre, amg=amg)
    x, p = _gpu_solve(A, P, solver, s.b, 1e-8)
    xs, ps = _gpu_solve(A, P, solver, s.b, 1e-11)
    rs = _rel_residual(s, xs)
    assert rs <= 2e-11, rs
    xo, po = oracle.matrix(s).solve(solver, pre, s.b, rel_tol=1e-8, max_iter=2000, mg=mg)
    h, ho = p["history"], po["history"]
    e_gpu = np.linalg.norm(x - xs) / np.linalg.norm(xs)
    e_ora = np.linalg.norm(xo - xs) / np.linalg.norm(xs)
    e_diff = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"{name}: iterations GPU {p['nIterations']} oracle {po['nIterations']} "
          f"(delta {p['nIterations'] - po['nIterations']}); error vs x*: GPU {e_gpu:.2e} oracle {e_ora:.2e}; "
          f"GPU vs oracle {e_diff:.2e}; x* true rel. residual {rs:.1e}")
    assert abs(h[0] - ho[0]) <= 1e-11 * ho[0]
    assert p["converged"] and po["converged"]
    assert abs(p["nIterations"] - po["nIterations"]) <= max(1, po["nIterations"] // 10)
    # CG: the first quarter of the history; BiCGStab: its first 20 iterations
    q = min(21, len(h), len(ho)) if solver == "PBiCGStab" else max(2, min(len(h), len(ho)) // 4)
    d = np.abs(h[: min(len(h), len(ho))] - ho[: min(len(h), len(ho))]) / ho[: min(len(h), len(ho))]
    print(f"{name}: history within 1e-6 up to iteration {int(np.argmax(d > 1e-6)) if np.any(d > 1e-6) else len(d)}"
          f" of {po['nIterations']}")
    assert np.all(d[:q] <= 1e-6)
    assert _rel_residual(s, x) <= 2e-8
    assert e_gpu <= 10 * e_ora + 1e-12, (e_gpu, e_ora)


@pytest.fixture
def reference_order():
    hip.set_reduction_order(hip.REDUCE_REFERENCE)
    yield
    hip.set_reduction_order(hip.REDUCE_TREE)


REF_CASES = [
    ("C1-type PCG+DIC 64^3", 1, "PCG", "DIC", None, 1e-6),
    ("C2 PCG+AMG(GS) 64^3", 2, "PCG", "AMG", [1, 500, 1, 0, 2, 2], 1e-8),
    ("C3 PBiCGStab+DILU 64^3", 3, "PBiCGStab", "DILU", None, 1e-8),
    ("C4 FP64 drop-in PCG+AMG(Cheby) 64^3", 4, "PCG", "AMG", [2, 500, 1, 0, 4, 4], 1e-8),
]


@pytest.mark.parametrize("name,c,solver,pre,mg,tol", REF_CASES, ids=[t[0] for t in REF_CASES])
def test_config_64_reference_order_bitwise(oracle, reference_order, name, c, solver, pre, mg, tol):
    s = meshgen.config_system(c, 64)
    A = hip.LduMatrix.from_system(s)
    amg = None
    if mg is not None:
        amg = hip.amg_params(smoother=mg[0], n_cells_coarsest=mg[1], n_vcycles=mg[2], n_pre_sweeps=mg[3],
                             n_post_sweeps=mg[4], n_finest_sweeps=mg[5])
    P = hip.Precond(A, pre, amg=amg)
    x, p = _gpu_solve(A, P, solver, s.b, tol, max_iter=1000)
    xo, po = oracle.matrix(s).solve(solver, pre, s.b, rel_tol=tol, max_iter=1000, mg=mg)
    print(f"{name}: {p['nIterations']} iterations (oracle {po['nIterations']})")
    assert p["nIterations"] == po["nIterations"]
    assert np.array_equal(p["history"], po["history"])
    assert np.array_equal(x, xo)
    assert p["converged"] and po["converged"]


def test_c4_fp64_drop_in_256():
    """The reference algorithm in FP64 at 256^3: PCG around the V-cycle over
    the full pairwise hierarchy (Chebyshev(4) smoothers, nPost + level sweeps,
    coarsest level 500 by PCG+DIC to 1e-2)."""
    s = meshgen.config_system(4, 256)
    A = hip.LduMatrix.from_system(s)
    P = hip.Precond(A, "AMG", amg=hip.amg_params(smoother=2, n_cells_coarsest=500, n_vcycles=1, n_post_sweeps=4,
                                                   n_finest_sweeps=4))
    S = hip.Solver("PCG", P, rel_tol=1e-8, max_iter=1000)
    b = hip.DeviceArray.from_numpy(s.b)
    x = hip.DeviceArray(s.n)
    _, p1 = S.solve(A, b, x=x)
    x1 = x.numpy()
    x.zero()
    _, p2 = S.solve(A, b, x=x)
    assert np.array_equal(x.numpy(), x1) and p1["nIterations"] == p2["nIterations"]
    assert p1["converged"] and p1["nIterations"] < 200, p1
    r = _rel_residual(s, x1)
    print(f"C4 FP64 drop-in 256^3: {p1['nIterations']} iterations, true rel. residual {r:.2e}")
    assert r <= 2e-8
