"""COO / CSR -> LDU built on the device (setup_dev.hip: the matrix builders
unap_hip_ldu_create_coo / _csr and the Hypre reader use it for matrices of
4096 cells or more) against the host conversion (unap_hip_coo_to_ldu), which
is pinned to the reference in tests/test_formats.py: the same faces and
values bit for bit, and the same refusals with the same messages.
"""
import time

import numpy as np
import pytest

from unap_amd import hip, meshgen

pytestmark = pytest.mark.gpu


def _coo(s, seed, base=0):
    rows = np.concatenate([np.arange(s.n), s.l, s.u])
    cols = np.concatenate([np.arange(s.n), s.u, s.l])
    vals = np.concatenate([s.diag, s.upper, s.lower_or_upper()])
    o = np.random.default_rng(seed).permutation(rows.size)  # any entry order
    return rows[o] + base, cols[o] + base, vals[o]


def _same(A, ref):
    got = A.arrays()
    for k in (0, 1, 3, 4):
        assert np.array_equal(got[k], ref[k]), k
    if got[2] is not None:
        assert np.array_equal(got[2], ref[2])


@pytest.mark.parametrize("c,N,symm,base", [(2, 17, True, 0), (3, 17, False, 1), (2, 20, False, 0)])
def test_device_coo_matches_host(c, N, symm, base):
    s = meshgen.config_system(c, N)
    rows, cols, vals = _coo(s, c, base)
    print(f"device COO: C{c} {N}^3 symm={symm} base={base}", flush=True)
    ref = hip.coo_to_ldu(s.n, rows, cols, vals, symmetric=symm, base=base)
    _same(hip.LduMatrix.from_coo(s.n, rows, cols, vals, symmetric=symm, base=base), ref)
    # CSR: rows sorted, columns in any order within a row
    o = np.argsort(rows, kind="stable")
    rs = np.concatenate([[0], np.cumsum(np.bincount(rows[o] - base, minlength=s.n))]).astype(np.int64) + base
    B = hip.LduMatrix.from_csr(s.n, rs, cols[o], vals[o], symmetric=symm, base=base)
    _same(B, ref)


def _message(fn):
    with pytest.raises(hip.UnapHipError) as e:
        fn()
    return str(e.value)


def test_device_coo_refusals():
    s = meshgen.config_system(3, 17)
    rows, cols, vals = _coo(s, 5)
    cases = {}
    r, c, v = rows.copy(), cols.copy(), vals.copy()
    r[7] = s.n  # out of range
    cases["range"] = (r, c, v)
    cases["duplicate"] = (np.append(rows, rows[3]), np.append(cols, cols[3]), np.append(vals, 1.0))
    k = int(np.flatnonzero(cols < rows)[0])  # drop one lower entry: an upper entry has no lower one
    cases["no lower"] = (np.delete(rows, k), np.delete(cols, k), np.delete(vals, k))
    k = int(np.flatnonzero(cols > rows)[0])  # drop one upper entry: a lower entry has no upper one
    cases["no upper"] = (np.delete(rows, k), np.delete(cols, k), np.delete(vals, k))
    for name, (r, c, v) in cases.items():
        host = _message(lambda: hip.coo_to_ldu(s.n, r, c, v, symmetric=False))
        dev = _message(lambda: hip.LduMatrix.from_coo(s.n, r, c, v, symmetric=False))
        print(f"{name}: {dev}", flush=True)
        assert host.split(":")[-1] == dev.split(":")[-1], (name, host, dev)


def test_device_coo_large():
    s = meshgen.config_system(3, 128)
    rows, cols, vals = _coo(s, 9)
    t0 = time.perf_counter()
    ref = hip.coo_to_ldu(s.n, rows, cols, vals, symmetric=False)
    t1 = time.perf_counter()
    A = hip.LduMatrix.from_coo(s.n, rows, cols, vals, symmetric=False)
    t2 = time.perf_counter()
    print(f"COO -> LDU at 128^3 ({rows.size} entries): host {t1 - t0:.2f} s, device builder {t2 - t1:.2f} s",
          flush=True)
    _same(A, ref)
