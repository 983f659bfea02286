"""GPU parity on hand-built LDU systems at the edges of the layout and the
level schedule, against the CPU oracle:

  * one cell and no faces, and a 70-cell diagonal matrix (empty SELL slices,
    one schedule level, partial last slice and wavefront);
  * 700-cell chains, symmetric and asymmetric (700 forward and backward levels:
    the longest dependency chain a sweep can meet, all in the one-workgroup
    tail);
  * a chain cut into pieces with an isolated cell (disconnected components:
    several breadth-first seeds, rows with no entries between full ones).

Bars: Amul, preconditioner application and Gauss-Seidel bit-exact; solvers in
reference-order reductions bit-exact (iterations, residual history, solution);
the batched solve (FP32 preconditioners) to its tolerance and the FP64 error.
"""
import numpy as np
import pytest

from unap_amd import hip, meshgen

pytestmark = pytest.mark.gpu


def _chain(n, asym=False, cuts=(), seed=0):
    """Cells 0..n-1 joined i -- i+1 except where i is in `cuts`; faces in
    owner order.  Diagonally dominant, so every solve converges."""
    rng = np.random.default_rng(seed)
    l = np.array([i for i in range(n - 1) if i not in cuts], dtype=np.int64)
    u = l + 1
    nf = l.shape[0]
    upper = -1.0 - 0.5 * rng.random(nf)
    lower = (-0.4 - 0.5 * rng.random(nf)) if asym else None
    diag = 3.5 + rng.random(n)
    b = rng.uniform(-1, 1, n)
    return meshgen.LduSystem(n, l, u, lower, diag, upper, b, not asym)


def _diagonal(n, seed=0):
    rng = np.random.default_rng(seed)
    e = np.zeros(0, dtype=np.int64)
    return meshgen.LduSystem(n, e, e.copy(), None, 1.0 + rng.random(n), np.zeros(0), rng.uniform(-1, 1, n), True)


CASES = {
    "cell": lambda: _diagonal(1),
    "diagonal70": lambda: _diagonal(70, 1),
    "chain700": lambda: _chain(700, seed=2),
    "chain700_asym": lambda: _chain(700, asym=True, seed=3),
    "pieces600": lambda: _chain(600, cuts=(149, 150, 299, 450), seed=4),
    "pieces600_asym": lambda: _chain(600, asym=True, cuts=(149, 150, 299, 450), seed=5),
}


@pytest.fixture
def reference_order():
    hip.set_reduction_order(hip.REDUCE_REFERENCE)
    yield
    hip.set_reduction_order(hip.REDUCE_TREE)


@pytest.mark.parametrize("case", sorted(CASES))
def test_edge_amul_precond_gs(oracle, case):
    s = CASES[case]()
    A = hip.LduMatrix.from_system(s)
    O = oracle.matrix(s)
    info = A.info()
    assert info["n"] == s.n and info["nfaces"] == s.nfaces
    x = np.random.default_rng(11).uniform(-1, 1, s.n)
    assert np.array_equal(A.spmv(x), O.spmv(x)), case
    kinds = ["diag", "DILU"] + (["DIC"] if s.symmetric else [])
    for kind in kinds:
        P = hip.Precond(A, kind)
        w_ref, rd_ref = O.precond(kind, x)
        assert np.array_equal(P.rD(), rd_ref), (case, kind)
        assert np.array_equal(P.apply(x), w_ref), (case, kind)
    S = hip.Smoother("GS")
    for sweeps in (1, 2):
        assert np.array_equal(S.smooth(A, x, s.b, sweeps), O.smooth("GS", x, s.b, sweeps)), (case, sweeps)


@pytest.mark.parametrize("lu", [True, False])
@pytest.mark.parametrize("case", sorted(CASES))
def test_edge_krylov_reference_order(oracle, reference_order, case, lu):
    # DIC/DILU of a chain is its exact factorisation (no fill-in), so those
    # solves end after one iteration; the diagonal preconditioner iterates
    s = CASES[case]()
    A = hip.LduMatrix.from_system(s)
    solver = "PCG" if s.symmetric else "PBiCGStab"
    pre = ("DIC" if s.symmetric else "DILU") if lu else "diag"
    x, p = hip.Solver(solver, hip.Precond(A, pre), rel_tol=1e-10, max_iter=400).solve(A, s.b)
    xo, po = oracle.matrix(s).solve(solver, pre, s.b, rel_tol=1e-10, max_iter=400)
    assert p["nIterations"] == po["nIterations"], (case, pre, p["nIterations"], po["nIterations"])
    assert np.array_equal(p["history"], po["history"]), (case, pre)
    assert np.array_equal(x, xo), (case, pre)
    assert p["converged"] and po["converged"], (case, pre)
    if s.nfaces == 0 or lu:
        assert p["nIterations"] == 1, (case, pre)


@pytest.mark.parametrize("case", ["chain700", "pieces600"])
def test_edge_mg_reference_order(oracle, reference_order, case):
    # agglomeration of chains (pairs along the chain, singletons at the cuts)
    # and the V-cycle on it, bitwise
    s = CASES[case]()
    A = hip.LduMatrix.from_system(s)
    mg = hip.amg_params(smoother=1, n_cells_coarsest=20)
    x, p = hip.Solver("MG", mg=mg, rel_tol=1e-8, max_iter=60).solve(A, s.b)
    xo, po = oracle.matrix(s).solve("MG", None, s.b, rel_tol=1e-8, max_iter=60, mg=[1, 20, 1, 0, 2, 2])
    assert p["nIterations"] == po["nIterations"], (case, p["nIterations"], po["nIterations"])
    assert np.array_equal(p["history"], po["history"]), case
    assert np.array_equal(x, xo), case


def test_edge_batch(oracle):
    # the edge systems inside one batched launch, beside ordinary ones: FP32
    # DIC / DILU in FP64 flexible CG / BiCGStab, each system checked against
    # its own residual and the oracle's FP64 solution
    systems = [CASES[k]() for k in sorted(CASES)]
    specs = meshgen.batch_specs(nsys=3, lo=4, hi=8, seed=21)
    systems += [meshgen.batch_system(sp) for sp in specs]
    solvers = ["PCG" if s.symmetric else "PBiCGStab" for s in systems]
    B = hip.Batch(systems, solvers)
    b = np.concatenate([s.b for s in systems])
    x, perfs = B.solve(b, rel_tol=1e-8, max_iter=1000)
    off = B.cell_off
    for i, (s, kind) in enumerate(zip(systems, solvers)):
        xi = x[off[i]:off[i + 1]]
        assert perfs[i]["converged"], (i, kind, perfs[i])
        M = oracle.matrix(s)
        assert np.linalg.norm(s.b - M.spmv(xi)) <= 2e-8 * np.linalg.norm(s.b), (i, kind)
        pre = "DIC" if kind == "PCG" else "DILU"
        # error bound as for the other batches: within 100x the error of the
        # oracle's FP64 solve to the same tolerance, or 1e-6
        xo, _ = M.solve(kind, pre, s.b, rel_tol=1e-8, max_iter=1000)
        x_ref, _ = M.solve(kind, pre, s.b, rel_tol=1e-13, max_iter=5000)
        e_batch = np.linalg.norm(xi - x_ref) / np.linalg.norm(x_ref)
        e_fp64 = np.linalg.norm(xo - x_ref) / np.linalg.norm(x_ref)
        assert e_batch <= max(100 * e_fp64, 1e-6), (i, kind, e_batch, e_fp64)
