"""CPU tests of the product library: it loads, exports every symbol its headers
declare, refuses bad input with a status code, and its host-side setup (cell
ordering, level schedules, pairwise agglomeration, Galerkin coarse matrices)
matches the oracle and the compiled reference bit for bit.  No GPU compute
call is made here: matrix and agglomeration handles, maps and coarse
addressing build on the host (the Galerkin values need the GPU)."""
import os
import re

import numpy as np
import pytest

from unap_amd import hip, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(unap_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = hip.lib()
    names = _declared("unap_hip.h")
    assert len(names) >= 35
    for name in names:
        assert hasattr(L, name), name
        assert name in hip.SIGNATURES, name
    M = meshgen._load()
    for name in _declared("unap_meshgen.h"):
        assert hasattr(M, name), name


def test_bad_arguments_return_status():
    with pytest.raises(hip.UnapHipError) as e:
        hip.LduMatrix(4, [0, 1], [1, 7], None, np.ones(4), [-1.0, -1.0])
    assert e.value.code == -1
    with pytest.raises(hip.UnapHipError):
        hip.LduMatrix(0, [], [], None, np.ones(1), [])
    A = hip.LduMatrix(3, [0, 1], [1, 2], None, np.full(3, 2.0), [-1.0, -1.0])
    with pytest.raises(hip.UnapHipError):
        hip.Solver("PCG", max_iter=-1)
    with pytest.raises(hip.UnapHipError):
        hip.amg_params(smoother=5) and hip.Precond(A, "AMG", amg=hip.amg_params(smoother=5))


# 48^3 (325k faces) takes the multi-threaded row-face pass
@pytest.mark.parametrize("c,N", [(1, 8), (2, 9), (3, 9), (4, 8), (3, 48)])
def test_level_schedule_and_order(c, N):
    s = meshgen.config_system(c, N)
    A = hip.LduMatrix.from_system(s)
    info = A.info()
    assert info["n"] == s.n and info["nfaces"] == s.nfaces and info["symmetric"] == (c != 3)
    # the schedule depth equals the longest owner->neighbour chain
    lev = np.zeros(s.n, np.int64)
    for f in range(s.nfaces):  # faces are owner-sorted, so owners are final first
        lev[s.u[f]] = max(lev[s.u[f]], lev[s.l[f]] + 1)
    assert info["fwd_levels"] == lev.max() + 1
    blev = np.zeros(s.n, np.int64)
    for f in range(s.nfaces - 1, -1, -1):
        blev[s.l[f]] = max(blev[s.l[f]], blev[s.u[f]] + 1)
    assert info["bwd_levels"] == blev.max() + 1


# the last two are large enough (over 65536 faces) for the multi-threaded
# passes of the coarse addressing
@pytest.mark.parametrize("c,N,nc", [(2, 10, 20), (3, 10, 20), (2, 13, 50), (4, 12, 30), (1, 9, 10), (2, 48, 500),
                                    (3, 40, 200)])
def test_agglomeration_matches_oracle(oracle, c, N, nc):
    # maps and coarse addressing (host); the Galerkin values are formed on
    # the GPU and compared in tests/test_gpu_galerkin.py
    s = meshgen.config_system(c, N)
    A = hip.LduMatrix.from_system(s)
    w = np.abs(s.upper)
    got = hip.Agglomeration(A, w, n_coarsest=nc).levels(values=False)
    ref = oracle.matrix(s).agglomerate(w, n_coarsest=nc, matrices=False)
    assert len(got) == len(ref) > 0
    lref = oracle.matrix(s).agglomerate(w, n_coarsest=nc)
    for g, r, rl in zip(got, ref, lref):
        for k in ("restrict", "face_restrict"):
            assert np.array_equal(g[k], r[k]), k
        for k in ("l", "u"):
            assert np.array_equal(g[k], rl[k]), k
        assert g["symmetric"] == rl["symmetric"]


@pytest.mark.parametrize("c,N,seed", [(2, 11, 5), (3, 11, 6)])
def test_agglomeration_matches_reference(reference, c, N, seed):
    s = meshgen.config_system(c, N, seed=seed)
    A = hip.LduMatrix.from_system(s)
    w = np.abs(s.upper) / np.linalg.norm(s.upper)
    got = hip.Agglomeration(A, w, n_coarsest=25).levels(values=False)
    ref = reference.matrix(s).agglomerate(w, n_coarsest=25)
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        for k in ("restrict", "face_restrict", "l", "u"):
            assert np.array_equal(g[k], r[k]), k


def test_set_topology_host_side():
    # setMatrixTopology (lduMatrix.cpp:566-597): new addressing, same cells.
    # Host-side effects only (no GPU call): the level counts follow the new
    # faces, and another face count leaves the matrix refusing use until its
    # coefficients are set.
    s = meshgen.config_system(2, 6)
    A = hip.LduMatrix.from_system(s)
    chain_l = np.arange(s.n - 1, dtype=np.int64)
    A.set_topology(chain_l, chain_l + 1)
    info = A.info()
    assert info["nfaces"] == s.n - 1 and info["fwd_levels"] == s.n and info["bwd_levels"] == s.n
    with pytest.raises(hip.UnapHipError) as e:
        A.spmv(np.ones(s.n))
    assert e.value.code == -4
    with pytest.raises(hip.UnapHipError) as e:
        A.set_topology([0, 1], [1, s.n])
    assert e.value.code == -1


def test_reference_binding_header_compiles():
    """unap_amd/integration/hipSolvers.hpp (the binding a maintainer adds to the
    reference tree) compiles against the reference's own headers, with the
    reference's build flags (its SConstruct: -DSCALAR_FLOAT64 -DLABEL_INT64)."""
    import glob
    import subprocess
    ref = "/root/reference/src"
    if not os.path.isdir(ref):
        pytest.skip("the reference sources are not here")
    inc = []
    for d, _, _ in os.walk(ref):
        if "swArrays" not in d and "swAthread" not in d:
            inc += ["-I", d]
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-DSCALAR_FLOAT64", "-DLABEL_INT64", "-w", *inc,
           "-I", "/opt/conda/include", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "unap_amd", "integration"),
           os.path.join(ROOT, "tests", "cpp", "ref_binding_example.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
