"""The Fortran / C entry points (libunap_hip_fortran.so) on the GPU, called
from a Fortran program (tests/fortran/unap_fortran_driver.f90, built by
build() where amdflang is installed) and through ctypes, against the oracle.

Both run the library's sums in the reference's order, so every case must
reproduce the oracle bit for bit: iterations, final residual and solution.
The one-shot entry points must leave a zeroed matrix handle behind
(fortranInterface.cpp:267, 345, 463: they delete the matrix).
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from unap_amd import hip, meshgen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "unap_amd", "lib")
DRIVER = os.path.join(LIB, "unap_fortran_driver")
I64P = ctypes.POINTER(ctypes.c_int64)
F64P = ctypes.POINTER(ctypes.c_double)
MG_GS = [1, 50, 1, 0, 2, 2]  # Gauss-Seidel, 50 coarsest cells, nPre 0, nPost 2, nFinest 2
MG_CHEBY = [2, 50, 1, 0, 2, 2]


@pytest.fixture
def reference_order():
    hip.set_reduction_order(hip.REDUCE_REFERENCE)
    yield
    hip.set_reduction_order(hip.REDUCE_TREE)


def _coo1(s, seed):
    """1-based COO grouped by row, columns shuffled within each row: the form
    reorderCOO (matrixConversion.cpp:38-99) sorts."""
    rows = np.concatenate([np.arange(s.n), s.l, s.u])
    cols = np.concatenate([np.arange(s.n), s.u, s.l])
    vals = np.concatenate([s.diag, s.upper, s.lower_or_upper()])
    o = np.lexsort((np.random.default_rng(seed).random(rows.size), rows))
    return rows[o] + 1, cols[o] + 1, vals[o]


def _set2(s):
    f = np.arange(s.nfaces)
    scale = 0.5 + 0.5 * np.abs(np.sin(0.37 * f))
    up2 = s.upper * scale
    lo2 = None if s.symmetric else s.lower * scale[::-1]
    # keep the diagonal dominance of the first set
    off = np.zeros(s.n)
    np.add.at(off, s.l, np.abs(up2))
    np.add.at(off, s.u, np.abs(up2 if lo2 is None else lo2))
    d2 = np.maximum(s.diag, 1.01 * off)
    return d2, up2, lo2


def _weights(s):
    return np.random.default_rng(s.n).uniform(0.1, 1.0, s.nfaces)


def _run_driver(tmp_path, s, tag):
    rows, cols, vals = _coo1(s, 3)
    d2, up2, lo2 = _set2(s)
    w = _weights(s)
    inp, out = tmp_path / f"{tag}.bin", tmp_path / f"{tag}.out"
    with open(inp, "wb") as f:
        np.array([s.n, rows.size, s.nfaces, int(s.symmetric)], np.int64).tofile(f)
        for a in (rows, cols):
            np.asarray(a, np.int64).tofile(f)
        for a in (vals, s.b):
            np.asarray(a, np.float64).tofile(f)
        for a in (s.l, s.u):
            np.asarray(a, np.int64).tofile(f)
        for a in (s.diag, s.upper) + (() if s.symmetric else (s.lower,)) + (d2, up2) + \
                 (() if s.symmetric else (lo2,)) + (w,):
            np.asarray(a, np.float64).tofile(f)
    print(f"fortran driver: {tag} n={s.n}", flush=True)
    r = subprocess.run([DRIVER, str(inp), str(out), "reference"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "unap_fortran_driver: done" in r.stdout
    raw = np.fromfile(out, np.uint8)
    rec = 8 * (3 + s.n)
    assert raw.size % rec == 0
    cases = []
    for k in range(raw.size // rec):
        chunk = raw[k * rec:(k + 1) * rec]
        it, handle = chunk[:16].view(np.int64)
        res = chunk[16:24].view(np.float64)[0]
        cases.append((int(it), int(handle), float(res), chunk[24:].view(np.float64).copy()))
    return cases, (d2, up2, lo2, w)


def _same(case, ref, what, one_shot=True):
    it, handle, res, x = case
    xo, po = ref
    if one_shot:
        assert handle == 0, (what, "the one-shot solvers consume the matrix")
        assert it == po["nIterations"], (what, it, po["nIterations"])
    assert res == po["final"], (what, res, po["final"])
    assert np.array_equal(x, xo), what


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="no Fortran compiler: driver not built")
def test_fortran_program_symmetric(oracle, tmp_path):
    s = meshgen.config_system(2, 14)
    cases, (d2, up2, _, w) = _run_driver(tmp_path, s, "sym")
    assert len(cases) == 7
    M = oracle.matrix(s)
    kw = dict(rel_tol=1e-8, min_iter=1)
    _same(cases[0], M.solve("PCG", "DIC", s.b, **kw), "coo + PCG/DIC")
    _same(cases[1], M.solve("PCG", "diag", s.b, **kw), "csr + PCG/diag")
    _same(cases[2], M.solve("PCG", "DIC", s.b, **kw), "ldu + PCG/DIC")
    _same(cases[3], M.solve("MG", None, s.b, max_iter=100, mg=MG_GS, **kw), "MG |upper| GS")
    _same(cases[4], M.solve("MG", None, s.b, max_iter=100, mg=MG_CHEBY, weights=w, **kw), "MG faceArea Chebyshev")
    # persistent: minIter stays at the solver default 0
    P = M.persistent("MG", mg=MG_GS, weights=w)
    _same(cases[5], P.solve(s.b, max_iter=100, rel_tol=1e-8), "sw MG", one_shot=False)
    M.set_coeffs(d2, up2)
    _same(cases[6], P.solve(s.b, max_iter=100, rel_tol=1e-8), "sw MG, set 2", one_shot=False)
    assert cases[5][1] != 0  # the persistent matrix survives its solves


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="no Fortran compiler: driver not built")
def test_fortran_program_asymmetric(oracle, tmp_path):
    s = meshgen.config_system(3, 12)
    cases, (d2, up2, lo2, _) = _run_driver(tmp_path, s, "asym")
    assert len(cases) == 4
    M = oracle.matrix(s)
    kw = dict(rel_tol=1e-8, min_iter=1)
    _same(cases[0], M.solve("PBiCGStab", "DILU", s.b, **kw), "coo + PBiCGStab/DILU")
    _same(cases[1], M.solve("PBiCGStab", "DILU", s.b, **kw), "ldu + PBiCGStab/DILU")
    P = M.persistent("PBiCGStab", "DILU")
    _same(cases[2], P.solve(s.b, min_iter=1, rel_tol=1e-8), "sw PBiCGStab", one_shot=False)
    # the DILU factor stays the one built at set_precond (lduDILUPrecond keeps rD)
    M.set_coeffs(d2, up2, lo2)
    _same(cases[3], P.solve(s.b, min_iter=1, rel_tol=1e-8), "sw PBiCGStab, set 2", one_shot=False)


def _flib():
    return ctypes.CDLL(os.path.join(LIB, "libunap_hip_fortran.so"))


def test_ldumatrixcreat_aliasing_flag(oracle, reference_order):
    """ldumatrixcreat_ with lower != upper: the asymmetric matrix by default;
    with the reference's aliasing, the symmetric matrix of the lower values."""
    L = _flib()
    s = meshgen.config_system(3, 10)
    n, nf = ctypes.c_int64(s.n), ctypes.c_int64(s.nfaces)
    l = np.ascontiguousarray(s.l, np.int64)
    u = np.ascontiguousarray(s.u, np.int64)
    lo, d, up = (np.ascontiguousarray(a, np.float64) for a in (s.lower, s.diag, s.upper))
    x = np.random.default_rng(2).uniform(-1, 1, s.n)
    for alias in (0, 1):
        L.unap_hip_fortran_set_reference_aliasing(alias)
        h = ctypes.c_long(0)
        L.ldumatrixcreat_(ctypes.byref(h), ctypes.byref(n), ctypes.byref(nf), l.ctypes.data_as(I64P),
                          u.ctypes.data_as(I64P), lo.ctypes.data_as(F64P), d.ctypes.data_as(F64P),
                          up.ctypes.data_as(F64P))
        assert L.unap_hip_fortran_last_status() == 0 and h.value != 0
        A = hip.LduMatrix._from_handle(h.value)
        want = meshgen.LduSystem(s.n, s.l, s.u, None, s.diag, s.lower, None, True) if alias else s
        assert np.array_equal(A.spmv(x), oracle.matrix(want).spmv(x)), alias
        del A
    L.unap_hip_fortran_set_reference_aliasing(0)
