"""Solver iterations replayed from a HIP graph against the same solves
launched kernel by kernel: bit-identical iterations, residual history and
solution, for every solver and preconditioner kind.  Each solver solves
twice (the second solve replays the graph the first one captured, with the
program's own work and synchronisations in between) and again after a
coefficient update (which captures anew)."""
import numpy as np
import pytest

from unap_amd import hip, meshgen

pytestmark = pytest.mark.gpu


def _amg(**kw):
    return hip.amg_params(n_cells_coarsest=20, **kw)


CASES = [
    ("PCG", "DIC", None, 1, 16),
    ("PCG", "AMG", dict(smoother=1, n_vcycles=1), 2, 16),
    ("PCG", "AMG", dict(smoother=2, n_vcycles=2), 2, 16),
    ("PBiCGStab", "DILU", None, 3, 16),
    ("FBiCGStab", "DILU", None, 2, 12),
    ("FCG", "AMG", dict(smoother=2, n_vcycles=1, n_post_sweeps=4, n_finest_sweeps=4, precision=32,
                        pair_steps=2, cycle=1), 4, 16),
    ("PBiCGStab", "DILU32", None, 3, 16),
    ("MG", None, dict(smoother=1), 2, 16),
]


def _run(solver, pre, amg, c, N, graphs):
    hip.set_graph_replay(graphs)
    try:
        s = meshgen.config_system(c, N)
        A = hip.LduMatrix.from_system(s)
        if solver == "MG":
            S = hip.Solver("MG", mg=_amg(**amg), rel_tol=1e-8, max_iter=100)
        else:
            if pre == "DILU32":
                P = hip.Precond(A, "DILU", precision=32)
            else:
                P = hip.Precond(A, pre, amg=None if amg is None else _amg(**amg))
            S = hip.Solver(solver, P, rel_tol=1e-8, max_iter=500)
        out = [S.solve(A, s.b), S.solve(A, s.b)]
        rng = np.random.default_rng(2)
        up = s.upper * rng.uniform(0.5, 1.0, s.nfaces)
        lo = None if s.symmetric else s.lower * rng.uniform(0.5, 1.0, s.nfaces)
        A.set_coeffs(s.diag, up, lo)
        out.append(S.solve(A, s.b))
        return out
    finally:
        hip.set_graph_replay(True)


@pytest.mark.parametrize("solver,pre,amg,c,N", CASES, ids=[f"{c[0]}-{c[1]}-{c[3]}" for c in CASES])
def test_graph_replay_matches_launches(solver, pre, amg, c, N):
    eager = _run(solver, pre, amg, c, N, False)
    graph = _run(solver, pre, amg, c, N, True)
    for k, ((xe, pe), (xg, pg)) in enumerate(zip(eager, graph)):
        assert pe["nIterations"] > 2, (k, pe["nIterations"])
        assert pg["nIterations"] == pe["nIterations"], k
        assert np.array_equal(pg["history"], pe["history"]), k
        assert np.array_equal(xg, xe), k
