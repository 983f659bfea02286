"""Reverse Cuthill-McKee reordering (src/tools/rcm/cooRCM.cpp) and the LDU
renumbering of rcmLDU_rewrite, host code (no GPU).

Pinned against the compiled reference's rcmLDU_nowrite where its order is
defined (its neighbour sort uses a comparator that is no ordering, harmless
only while the neighbours met at each step share one degree: paths and
rings, in scrambled labels and face orientations), and by what the
algorithm fixes on any graph: a permutation whose reverse is a breadth-first
visit from a least-degree cell (distance from the seed never increases along
the new labels), and a renumbered matrix that is the same operator
(A'[new i, new j] = A[i, j]) with faces turned to owner < neighbour and
sorted.
"""
import ctypes

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import shortest_path

from unap_amd import hip, meshgen

I64P = ctypes.POINTER(ctypes.c_int64)


def _scrambled(n, edges, seed):
    """Faces of a graph under a random relabelling, random orientation and
    random face order."""
    rng = np.random.default_rng(seed)
    lab = rng.permutation(n)
    e = np.array([(lab[a], lab[b]) for a, b in edges], np.int64)
    flip = rng.random(len(e)) < 0.5
    e[flip] = e[flip][:, ::-1]
    e = e[rng.permutation(len(e))]
    return np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1])


@pytest.mark.parametrize("kind,n,seed", [("path", 12, 1), ("path", 30, 2), ("ring", 10, 3), ("ring", 25, 4)])
def test_rcm_matches_reference(reference, kind, n, seed):
    edges = [(i, i + 1) for i in range(n - 1)] + ([(n - 1, 0)] if kind == "ring" else [])
    l, u = _scrambled(n, edges, seed)
    nf = l.size
    ref_cells, ref_faces = np.empty(n, np.int64), np.empty(nf, np.int64)
    reference.tool("rcm_ldu")(nf, n, l.ctypes.data_as(I64P), u.ctypes.data_as(I64P),
                              ref_cells.ctypes.data_as(I64P), ref_faces.ctypes.data_as(I64P))
    mine = hip.rcm_order(n, l, u)
    assert np.array_equal(mine, ref_cells)
    _, _, fmap, _, _, _ = hip.renumber(n, l, u, mine)
    assert np.array_equal(fmap, ref_faces)


def _check_bfs_reverse(n, l, u, new_of_old):
    assert np.array_equal(np.sort(new_of_old), np.arange(n))
    old_of_new = np.empty(n, np.int64)
    old_of_new[new_of_old] = np.arange(n)
    deg = np.bincount(l, minlength=n) + np.bincount(u, minlength=n)
    seed = old_of_new[-1]  # the Cuthill-McKee sequence starts here
    assert deg[seed] == deg.min() and seed == np.flatnonzero(deg == deg.min())[0]
    g = coo_matrix((np.ones(l.size), (l, u)), shape=(n, n))
    dist = shortest_path(g, directed=False, unweighted=True, indices=seed)
    assert np.all(np.isfinite(dist))
    d = dist[old_of_new]
    assert np.all(np.diff(d) <= 0)  # reversed breadth-first: distance never grows


@pytest.mark.parametrize("c,N", [(2, 6), (3, 7), (1, 9)])
def test_rcm_is_reversed_bfs(c, N):
    s = meshgen.config_system(c, N)
    p = hip.rcm_order(s.n, s.l, s.u)
    _check_bfs_reverse(s.n, s.l, s.u, p)
    l, u, *_ = hip.renumber(s.n, s.l, s.u, p)
    # bandwidth: a reversed breadth-first order from a corner of an N^3 box
    assert np.max(u - l) < np.max(np.abs(s.u - s.l))
    assert np.max(u - l) <= 2 * N * N


@pytest.mark.parametrize("c,N", [(2, 5), (3, 5)])
def test_renumber_same_operator(c, N):
    s = meshgen.config_system(c, N)
    p = np.random.default_rng(9).permutation(s.n).astype(np.int64)
    lo = None if s.symmetric else s.lower
    l, u, fmap, lo2, d2, up2 = hip.renumber(s.n, s.l, s.u, p, s.diag, s.upper, lo)
    assert np.all(l < u)
    key = l * s.n + u
    assert np.all(np.diff(key) >= 0)
    A = s.dense()
    B = np.zeros((s.n, s.n))
    B[np.arange(s.n), np.arange(s.n)] = d2
    B[l, u] = up2
    B[u, l] = up2 if lo2 is None else lo2
    assert np.array_equal(B[np.ix_(p, p)], A)
    # the face map: old face f sits at k (or ~k when it turned)
    for f in range(s.nfaces):
        k = fmap[f]
        kk = k if k >= 0 else ~k
        a, b = p[s.l[f]], p[s.u[f]]
        assert (l[kk], u[kk]) == ((a, b) if k >= 0 else (b, a))


def test_renumber_refusals():
    with pytest.raises(hip.UnapHipError):
        hip.renumber(3, [0, 1], [1, 2], [0, 0, 1])  # not a permutation
    with pytest.raises(hip.UnapHipError):
        hip.rcm_order(3, [0, 1], [1, 3])
