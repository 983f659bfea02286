"""Argument checks of the batched solve's create call (host code, no GPU):
the refusals happen before the library touches a device, so they run here.

unap_hip_batch_create packs every system at the offsets the caller gives
into one device array indexed with 32-bit integers; offsets that do not
start at 0, that go backwards, or whose total reaches 2^31 would index
outside what it allocates, and are refused with UNAP_HIP_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

from unap_amd import hip

ERR_ARG = -1


def _create(cell_off, face_off, nfaces_total=None, ncells_total=None):
    cell_off = np.asarray(cell_off, np.int64)
    face_off = np.asarray(face_off, np.int64)
    nsys = len(cell_off) - 1
    # the face and cell arrays are sized to what a well-formed call with
    # these offsets would read; the refusals come before any of them is read
    nf = int(nfaces_total if nfaces_total is not None else max(face_off.max(), 1))
    nc = int(ncells_total if ncells_total is not None else max(cell_off.max(), 1))
    la = np.zeros(nf, np.int64)
    ua = np.ones(nf, np.int64)
    up = np.full(nf, -1.0)
    d = np.full(nc, 4.0)
    kinds = np.zeros(nsys, np.int32)
    h = ctypes.c_void_p()
    rc = hip.lib().unap_hip_batch_create(nsys, hip._p(cell_off, hip._i64p), hip._p(face_off, hip._i64p),
                                         hip._p(la, hip._i64p), hip._p(ua, hip._i64p), hip._p(up), hip._p(d),
                                         hip._p(up), hip._p(kinds, hip._i32p), ctypes.byref(h))
    return rc, hip.lib().unap_hip_last_error().decode(), h


@pytest.mark.parametrize("cell_off,face_off,what", [
    ([1, 3, 6], [0, 1, 2], "cell_off[0]"),
    ([0, 3, 6], [2, 3, 4], "face_off[0]"),
    ([0, 3, 3], [0, 1, 2], "increase"),          # an empty system
    ([0, 5, 3], [0, 1, 2], "increase"),          # cells going backwards
    ([0, 3, 6], [0, 2, 1], "increase"),          # faces going backwards
])
def test_batch_offsets_refused(cell_off, face_off, what):
    rc, msg, h = _create(cell_off, face_off)
    assert rc == ERR_ARG, (rc, msg)
    assert what in msg, msg
    assert not h.value


def test_batch_total_beyond_int32_refused():
    # offsets whose last entry reaches 2^31: refused from the offsets alone,
    # before any array is read (the arrays passed are small on purpose)
    rc, msg, h = _create([0, 3, 1 << 31], [0, 1, 2], ncells_total=8)
    assert rc == ERR_ARG, (rc, msg)
    assert "2^31" in msg, msg
    rc, msg, h = _create([0, 3, 6], [0, 1, 1 << 31], nfaces_total=8)
    assert rc == ERR_ARG, (rc, msg)
    assert "2^31" in msg, msg


def test_batch_null_arguments_refused():
    h = ctypes.c_void_p()
    rc = hip.lib().unap_hip_batch_create(0, None, None, None, None, None, None, None, None, ctypes.byref(h))
    assert rc == ERR_ARG
