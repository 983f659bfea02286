"""The setup built on the device (setup_dev.hip: cell orders, sweep
schedules, SELL-32 layouts in every order mode, the agglomeration's pairing
lists) against the host builder, field for field, through
unap_hip_setup_self_check.

Cases: the workload laws at odd and even sizes, a matrix whose breadth-first
visit has many layers and whose sweeps have many levels (reverse
Cuthill-McKee numbering), the test/ex2.cpp ring (faces not owner-sorted: no
schedules), and a graph of several components with isolated cells.
"""
import numpy as np
import pytest

from unap_amd import hip, meshgen

pytestmark = pytest.mark.gpu


def _renumbered(s, new_of_old):
    lo = None if s.symmetric else s.lower
    l, u, _, lo2, d2, up2 = hip.renumber(s.n, s.l, s.u, new_of_old, s.diag, s.upper, lo)
    b = np.empty(s.n)
    b[new_of_old] = s.b
    return meshgen.LduSystem(s.n, l, u, lo2, d2, up2, b, s.symmetric)


def _cases():
    out = [("C1 12^3", meshgen.config_system(1, 12)), ("C2 17^3", meshgen.config_system(2, 17)),
           ("C3 16^3", meshgen.config_system(3, 16)), ("C2 33^3", meshgen.config_system(2, 33))]
    s = meshgen.config_system(2, 20)
    out.append(("C2 20^3 RCM", _renumbered(s, hip.rcm_order(s.n, s.l, s.u))))
    n = 10  # test/ex2.cpp ring: the last face (0, 9) breaks owner order
    l = np.array([i if i < n - 1 else 0 for i in range(n)])
    u = np.array([i + 1 if i < n - 1 else n - 1 for i in range(n)])
    out.append(("ex2 ring", meshgen.LduSystem(n, l, u, None, np.full(n, 4.0), np.full(n, -1.0), None, True)))
    # two boxes and isolated cells, interleaved labels
    a, b = meshgen.config_system(2, 6), meshgen.config_system(2, 7)
    n = a.n + b.n + 5
    lab = np.random.default_rng(4).permutation(n)
    la = np.concatenate([lab[a.l], lab[a.n + b.l]])
    ua = np.concatenate([lab[a.u], lab[a.n + b.u]])
    lo, hi = np.minimum(la, ua), np.maximum(la, ua)
    o = np.lexsort((hi, lo))
    up = np.concatenate([a.upper, b.upper])[o]
    d = np.full(n, 10.0)
    out.append(("two boxes + isolated", meshgen.LduSystem(n, lo[o], hi[o], None, d, up, None, True)))
    return out


@pytest.mark.parametrize("name,s", _cases(), ids=[c[0] for c in _cases()])
def test_device_setup_matches_host(name, s):
    print(f"setup self-check: {name} n={s.n}", flush=True)
    A = hip.LduMatrix.from_system(s)
    assert hip.setup_self_check(A) == 0, name
    w = np.random.default_rng(1).uniform(0.1, 1.0, s.nfaces)
    assert hip.setup_self_check(A, w) == 0, name
