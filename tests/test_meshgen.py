"""The seeded box generator shared by the GPU path, the oracle and the bench."""
import ctypes

import numpy as np
import pytest

from unap_amd import meshgen


@pytest.mark.parametrize("c", [1, 2, 3, 4])
def test_structure(c):
    s = meshgen.config_system(c, 7)
    n, nf = s.n, s.nfaces
    assert n == 343 and nf == 3 * 49 * 6
    assert np.all(s.l < s.u)
    key = s.l * n + s.u
    assert np.all(np.diff(key) > 0)  # sorted by (owner, neighbour), no duplicates
    deg = np.bincount(s.l, minlength=n) + np.bincount(s.u, minlength=n)
    assert deg.min() == 3 and deg.max() == 6  # corners .. interior of a 7-point box


def test_permutation_and_grid_edges():
    N = 6
    s = meshgen.box_system(N, N, N, meshgen.POISSON_CONST, shift=1e-6, seed=1234, with_perm=True)
    p = s.perm
    assert np.array_equal(np.sort(p), np.arange(s.n))
    inv = np.empty_like(p)
    inv[p] = np.arange(s.n)
    a, b = inv[s.l], inv[s.u]  # original ids
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    d = hi - lo
    assert set(np.unique(d)) <= {1, N, N * N}
    ia, ja, ka = lo % N, (lo // N) % N, lo // (N * N)
    assert np.all((d != 1) | (ia < N - 1)) and np.all((d != N) | (ja < N - 1)) and np.all((d != N * N) | (ka < N - 1))


def test_coefficients():
    s = meshgen.config_system(1, 6)
    assert np.all(s.upper == -1.0) and np.all(s.diag == 6.0 + 1e-6)
    s = meshgen.config_system(2, 6)
    w = -s.upper
    assert np.all((w >= 1e-3) & (w <= 1.0))
    rs = np.zeros(s.n)
    np.add.at(rs, s.l, w)
    np.add.at(rs, s.u, w)
    assert np.allclose(s.diag - rs, 1e-6, rtol=0, atol=1e-12)
    A = s.dense()
    assert np.array_equal(A, A.T)
    s = meshgen.config_system(3, 6)
    # diffusion 1 plus upwind: off-diagonals <= -1, each face's pair sums to -(2 + |F|)
    assert np.all(s.upper <= -1.0) and np.all(s.lower <= -1.0)
    F = np.abs(s.upper - s.lower)
    assert np.all(F <= 5.0)
    colsum = np.zeros(s.n)  # A^T 1: diag minus the outflow = 1 + inflow per face ... column sums
    np.add.at(colsum, s.u, s.upper)
    np.add.at(colsum, s.l, s.lower)
    # row sums: diag_i + sum_j A(i,j) = sum_j (1 + out_ij) - (1 + in_ij) + 1e-3 = net outflow + 1e-3
    A = s.dense()
    assert np.all(np.abs(A.sum(axis=1)) <= 5 * 6 + 1e-3 + 1e-9)


def test_determinism_and_seeds():
    a = meshgen.config_system(2, 5)
    b = meshgen.config_system(2, 5)
    c = meshgen.config_system(2, 5, seed=77)
    assert np.array_equal(a.upper, b.upper) and np.array_equal(a.b, b.b) and np.array_equal(a.l, b.l)
    assert not np.array_equal(a.l, c.l)


def test_splitmix_twin():
    lib = meshgen._load()
    st = ctypes.c_uint64(12345)
    py = meshgen.SplitMix64(12345)
    for _ in range(50):
        assert lib.unap_meshgen_splitmix64(ctypes.byref(st)) == py.next()


def test_batch_specs():
    specs = meshgen.batch_specs()
    assert len(specs) == 512
    dims = np.array([sp[:3] for sp in specs])
    assert dims.min() >= 8 and dims.max() <= 48
    kinds = [sp[3] for sp in specs]
    assert kinds.count(meshgen.POISSON_LOGUNIF) == round(0.6 * 512)
    assert specs == meshgen.batch_specs()
