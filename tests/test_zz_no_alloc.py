"""The allocation-free Amul entry point (unap_hip_ldu_spmv_noalloc) against
unap_hip_ldu_spmv through the C ABI: bit-identical results on two calls in a
row, the second on a smaller matrix (the shared scratch keeps the first
call's contents), and the documented refusals."""
import numpy as np
import pytest

from unap_amd import hip, meshgen

pytestmark = pytest.mark.gpu


def _case(c, N, seed):
    s = meshgen.config_system(c, N)
    A = hip.LduMatrix.from_system(s)
    x = np.random.default_rng(seed).uniform(-1, 1, s.n)
    xd = hip.DeviceArray.from_numpy(x)
    ref = hip.DeviceArray(s.n)
    A.spmv(xd, out=ref)  # the existing entry point (also builds the device layout)
    return s, A, xd, ref


def test_noalloc_spmv_matches_spmv():
    big = _case(3, 40, 1)     # 64000 cells, asymmetric
    small = _case(2, 17, 2)   # 4913 cells, symmetric
    for s, A, xd, ref in (big, small):
        out = hip.DeviceArray(s.n)
        A.spmv_noalloc(xd, out)
        assert np.array_equal(out.numpy(), ref.numpy()), s.n
    # and back to the larger one after the smaller call
    s, A, xd, ref = big
    out = hip.DeviceArray(s.n)
    A.spmv_noalloc(xd, out)
    assert np.array_equal(out.numpy(), ref.numpy())


def test_noalloc_spmv_refusals():
    # a matrix whose device layout is not built yet: the call would allocate
    s = meshgen.config_system(2, 8)
    A = hip.LduMatrix.from_system(s)
    x = hip.DeviceArray.from_numpy(np.ones(s.n))
    y = hip.DeviceArray(s.n)
    with pytest.raises(hip.UnapHipError) as e:
        A.spmv_noalloc(x, y)
    assert e.value.code == -4
    # more cells than the scratch holds (a chain of 2^21 + 1 cells)
    n = (1 << 21) + 1
    l = np.arange(n - 1, dtype=np.int64)
    B = hip.LduMatrix(n, l, l + 1, None, np.full(n, 2.0), np.full(n - 1, -1.0))
    xb = hip.DeviceArray(n)
    yb = hip.DeviceArray(n)
    with pytest.raises(hip.UnapHipError) as e:
        B.spmv_noalloc(xb, yb)
    assert e.value.code == -1
