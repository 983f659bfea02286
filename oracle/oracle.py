"""ctypes bindings of the CPU oracle (liboracle.so, prefix uo_) and of the shim
over the compiled reference (_ref/libunap_ref.so, prefix ur_).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libunap_ref.so")

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

PRECOND = {"none": 0, "diag": 1, "DIC": 2, "DILU": 3, "AMG": 4, "GS1": 5}  # GS1: oracle only
SOLVER = {"PCG": 1, "PBiCGStab": 2, "FCG": 4, "FBiCGStab": 5}
SMOOTHER = {"GS": 1, "Cheby": 2}


def _p(a, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


class CpuSolvePath:
    """One implementation of the flat oracle interface (oracle_capi.h)."""

    def __init__(self, path, prefix):
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        P = prefix
        L = self.lib
        _sig(L, P + "ldu_create", _vp, [ctypes.c_int64, ctypes.c_int64, _i64p, _i64p, _f64p, _f64p, _f64p])
        _sig(L, P + "ldu_destroy", None, [_vp])
        _sig(L, P + "spmv", None, [_vp, _f64p, _f64p])
        for nm in ("owner_start", "losort", "losort_start"):
            _sig(L, P + nm, None, [_vp, _i64p])
        _sig(L, P + "precond_create", _vp, [_vp, ctypes.c_int])
        _sig(L, P + "precond_apply", None, [_vp, _f64p, _f64p, ctypes.c_int64])
        _sig(L, P + "precond_rd", None, [_vp, _f64p, ctypes.c_int64])
        _sig(L, P + "precond_destroy", None, [_vp])
        _sig(L, P + "smoother_create", _vp, [ctypes.c_int])
        _sig(L, P + "smoother_smooth", None, [_vp, _vp, _f64p, _f64p, ctypes.c_int64])
        _sig(L, P + "smoother_init", None, [_vp])
        _sig(L, P + "smoother_destroy", None, [_vp])
        _sig(L, P + "eigen_diag_pcg", ctypes.c_double, [_vp, _f64p, _f64p, ctypes.c_int64])
        _sig(L, P + "aggl_create", _vp, [_vp, _f64p, ctypes.c_int64, ctypes.c_int64])
        _sig(L, P + "aggl_nlevels", ctypes.c_int64, [_vp])
        _sig(L, P + "aggl_level_sizes", None, [_vp, ctypes.c_int64, _i64p])
        _sig(L, P + "aggl_restrict_addr", None, [_vp, ctypes.c_int64, _i64p])
        _sig(L, P + "aggl_face_restrict_addr", None, [_vp, ctypes.c_int64, _i64p])
        _sig(L, P + "aggl_build_matrices", None, [_vp])
        _sig(L, P + "aggl_coarse_matrix", None, [_vp, ctypes.c_int64, _i64p, _i64p, _f64p, _f64p, _f64p, _i64p])
        _sig(L, P + "aggl_destroy", None, [_vp])
        _sig(L, P + "krylov_solve", ctypes.c_int,
             [_vp, ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
              ctypes.c_double, _f64p, _f64p, _f64p, _f64p, ctypes.c_int64, _i64p])
        _sig(L, P + "mg_solve", ctypes.c_int,
             [_vp, _i64p, _f64p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
              _f64p, _f64p, _f64p, _f64p, ctypes.c_int64, _i64p])

        _sig(L, P + "ldu_set_coeffs", None, [_vp, _f64p, _f64p, _f64p])
        _sig(L, P + "solver_create", _vp, [_vp, ctypes.c_int, ctypes.c_int, _i64p, _f64p])
        _sig(L, P + "solver_run", ctypes.c_int,
             [_vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _f64p, _f64p, _f64p,
              _f64p, ctypes.c_int64, _i64p])
        _sig(L, P + "solver_destroy", None, [_vp])
        _sig(L, P + "timed_solve", ctypes.c_int,
             [_vp, ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, _f64p, _f64p,
              _f64p, _f64p])

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    # ---- formats (the compiled reference's tools only) -----------------
    def _tools(self):
        L, P = self.lib, self.prefix
        if getattr(self, "_tools_bound", False):
            return
        _sig(L, P + "ldu_get", None, [_vp, _i64p, _i64p, _i64p, _i64p, _i64p, _f64p, _f64p, _f64p])
        _sig(L, P + "coo2ldu", _vp, [ctypes.c_int64, ctypes.c_int64, _i64p, _i64p, _f64p, ctypes.c_int])
        _sig(L, P + "csr2ldu", _vp, [ctypes.c_int64, _i64p, _i64p, _f64p, ctypes.c_int])
        _sig(L, P + "reorder_coo", None, [_f64p, _i64p, _i64p, ctypes.c_int64, ctypes.c_int64])
        _sig(L, P + "reorder_uface", None, [_i64p, _i64p, ctypes.c_int64, ctypes.c_int64, _i64p])
        _sig(L, P + "reorder_lface", None, [_i64p, _i64p, ctypes.c_int64, ctypes.c_int64, _i64p])
        _sig(L, P + "reorder_value", None, [_f64p, _i64p, ctypes.c_int64])
        _sig(L, P + "write_openfoam", None, [_vp, ctypes.c_char_p])
        _sig(L, P + "write_vector_openfoam", None, [_f64p, ctypes.c_int64, ctypes.c_char_p])
        _sig(L, P + "read_openfoam", _vp, [ctypes.c_char_p])
        _sig(L, P + "read_vector_openfoam", None, [ctypes.c_char_p, ctypes.c_int64, _f64p])
        _sig(L, P + "read_hypre", _vp, [ctypes.c_char_p])
        _sig(L, P + "read_vector_hypre", None, [ctypes.c_char_p, ctypes.c_int64, _f64p])
        if P == "ur_":
            _sig(L, P + "rcm_ldu", None, [ctypes.c_int64, ctypes.c_int64, _i64p, _i64p, _i64p, _i64p])
        self._tools_bound = True

    def ldu_arrays(self, h):
        """(lowerAddr, upperAddr, lower, diag, upper, symmetric) of a reference lduMatrix handle."""
        self._tools()
        n, nf, sy = (np.zeros(1, np.int64) for _ in range(3))
        self.fn("ldu_get")(h, _p(n, _i64p), _p(nf, _i64p), _p(sy, _i64p), None, None, None, None, None)
        m = int(nf[0])
        l = np.empty(m, np.int64)
        u = np.empty(m, np.int64)
        lo = np.empty(m)
        d = np.empty(int(n[0]))
        up = np.empty(m)
        self.fn("ldu_get")(h, None, None, None, _p(l, _i64p), _p(u, _i64p), _p(lo), _p(d), _p(up))
        return l, u, lo, d, up, bool(sy[0])

    def tool(self, name):
        self._tools()
        return self.fn(name)

    # ---- matrix --------------------------------------------------------
    def matrix(self, sys):
        return CpuMatrix(self, sys)


class CpuMatrix:
    def __init__(self, path, sys):
        self.path = path
        self.sys = sys
        self._keep = (np.ascontiguousarray(sys.l, np.int64), np.ascontiguousarray(sys.u, np.int64),
                      None if sys.symmetric else np.ascontiguousarray(sys.lower),
                      np.ascontiguousarray(sys.diag), np.ascontiguousarray(sys.upper))
        l, u, lo, d, up = self._keep
        self.h = path.fn("ldu_create")(sys.n, sys.nfaces, _p(l, _i64p), _p(u, _i64p), _p(lo), _p(d), _p(up))

    def __del__(self):
        try:
            self.path.fn("ldu_destroy")(self.h)
        except Exception:
            pass

    @property
    def n(self):
        return self.sys.n

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n)
        self.path.fn("spmv")(self.h, _p(x), _p(y))
        return y

    def addressing(self, which):
        size = self.n + 1 if which in ("owner_start", "losort_start") else self.sys.nfaces
        out = np.empty(size, np.int64)
        self.path.fn(which)(self.h, _p(out, _i64p))
        return out

    def precond(self, kind, r):
        P = self.path.fn("precond_create")(self.h, PRECOND[kind])
        r = np.ascontiguousarray(r, np.float64)
        w = np.empty(self.n)
        self.path.fn("precond_apply")(P, _p(r), _p(w), self.n)
        rd = np.empty(self.n)
        if kind != "none":
            self.path.fn("precond_rd")(P, _p(rd), self.n)
        self.path.fn("precond_destroy")(P)
        return w, rd

    def precond_obj(self, kind):
        """A preconditioner built now and kept (apply(r) -> w)."""
        return CpuPrecond(self, kind)

    def smooth(self, kind, x, b, nsweeps, calls=1):
        S = self.path.fn("smoother_create")(SMOOTHER[kind])
        x = np.array(x, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        for _ in range(calls):
            self.path.fn("smoother_smooth")(S, self.h, _p(x), _p(b), nsweeps)
        self.path.fn("smoother_destroy")(S)
        return x

    def eigen(self, x, b, npcg=10):
        x = np.array(x, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        lam = self.path.fn("eigen_diag_pcg")(self.h, _p(x), _p(b), npcg)
        return lam, x

    def agglomerate(self, weights, max_levels=50, n_coarsest=50, matrices=True):
        w = np.ascontiguousarray(weights, np.float64)
        G = self.path.fn("aggl_create")(self.h, _p(w), max_levels, n_coarsest)
        if matrices:
            self.path.fn("aggl_build_matrices")(G)
        levels = []
        nprev, fprev = self.n, self.sys.nfaces
        for lev in range(self.path.fn("aggl_nlevels")(G)):
            sz = np.empty(2, np.int64)
            self.path.fn("aggl_level_sizes")(G, lev, _p(sz, _i64p))
            nc, nfc = int(sz[0]), int(sz[1])
            ra = np.empty(nprev, np.int64)
            fa = np.empty(fprev, np.int64)
            self.path.fn("aggl_restrict_addr")(G, lev, _p(ra, _i64p))
            self.path.fn("aggl_face_restrict_addr")(G, lev, _p(fa, _i64p))
            lv = {"n": nc, "nfaces": nfc, "restrict": ra, "face_restrict": fa}
            if matrices:
                l = np.empty(nfc, np.int64)
                u = np.empty(nfc, np.int64)
                lo = np.empty(nfc)
                up = np.empty(nfc)
                d = np.empty(nc)
                sy = np.empty(1, np.int64)
                self.path.fn("aggl_coarse_matrix")(G, lev, _p(l, _i64p), _p(u, _i64p), _p(lo), _p(d), _p(up),
                                                   _p(sy, _i64p))
                lv.update(l=l, u=u, lower=lo, diag=d, upper=up, symmetric=bool(sy[0]))
            levels.append(lv)
            nprev, fprev = nc, nfc
        self.path.fn("aggl_destroy")(G)
        return levels

    def set_coeffs(self, diag, upper, lower=None):
        """setMatrixCoeffients: new values, same addressing (lower None: symmetric)."""
        d = np.ascontiguousarray(diag, np.float64)
        up = np.ascontiguousarray(upper, np.float64)
        lo = None if lower is None else np.ascontiguousarray(lower, np.float64)
        self._vals = (d, up, lo)  # the oracle restatement copies; the reference shim copies too
        self.path.fn("ldu_set_coeffs")(self.h, _p(d), _p(up), _p(lo))

    def persistent(self, solver, precond=None, mg=None, weights=None):
        """A solver whose preconditioner / hierarchy is built now and kept
        (MG: agglomerated with the given face weights, default |upper| / ||upper||)."""
        return CpuPersistentSolver(self, solver, precond, mg, weights)

    def solve(self, solver, precond, b, x0=None, max_iter=1000, min_iter=0, rel_tol=1e-6, tol=0.0,
              mg=None, hist_cap=2048, weights=None):
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64)
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        perf = np.zeros(6)
        hist = np.zeros(hist_cap)
        hl = np.zeros(1, np.int64)
        mga = None if mg is None else np.asarray(mg, np.int64)
        if solver == "MG":
            rc = self.path.fn("mg_solve")(self.h, _p(mga, _i64p), _p(w), max_iter, min_iter, rel_tol, tol,
                                          _p(x), _p(b), _p(perf), _p(hist), hist_cap, _p(hl, _i64p))
        else:
            rc = self.path.fn("krylov_solve")(self.h, SOLVER[solver], PRECOND[precond], _p(mga, _i64p),
                                              max_iter, min_iter, rel_tol, tol, _p(x), _p(b), _p(perf),
                                              _p(hist), hist_cap, _p(hl, _i64p))
        if rc != 0:
            raise RuntimeError(f"{self.path.prefix}solve failed: {rc}")
        h = hist[: min(int(hl[0]), hist_cap)].copy()
        return x, {"initial": perf[0], "final": perf[1], "nIterations": int(perf[2]),
                   "converged": bool(perf[3]), "singular": bool(perf[4]), "history": h}


class CpuPrecond:
    def __init__(self, M, kind):
        self.M = M
        self.h = M.path.fn("precond_create")(M.h, PRECOND[kind])

    def __del__(self):
        try:
            self.M.path.fn("precond_destroy")(self.h)
        except Exception:
            pass

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        w = np.empty(self.M.n)
        self.M.path.fn("precond_apply")(self.h, _p(r), _p(w), self.M.n)
        return w


class CpuPersistentSolver:
    def __init__(self, M, solver, precond, mg, weights=None):
        self.M = M
        self.mga = None if mg is None else np.asarray(mg, np.int64)
        self.w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        kind = 3 if solver == "MG" else SOLVER[solver]
        pre = 0 if precond is None else PRECOND[precond]
        self.h = M.path.fn("solver_create")(M.h, kind, pre, _p(self.mga, _i64p), _p(self.w))
        if not self.h:
            raise RuntimeError("solver_create failed")

    def __del__(self):
        try:
            self.M.path.fn("solver_destroy")(self.h)
        except Exception:
            pass

    def solve(self, b, x0=None, max_iter=1000, min_iter=0, rel_tol=1e-6, tol=0.0, hist_cap=2048):
        M = self.M
        x = np.zeros(M.n) if x0 is None else np.array(x0, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        perf = np.zeros(6)
        hist = np.zeros(hist_cap)
        hl = np.zeros(1, np.int64)
        rc = M.path.fn("solver_run")(self.h, M.h, max_iter, min_iter, rel_tol, tol, _p(x), _p(b), _p(perf),
                                     _p(hist), hist_cap, _p(hl, _i64p))
        if rc != 0:
            raise RuntimeError(f"solver_run failed: {rc}")
        h = hist[: min(int(hl[0]), hist_cap)].copy()
        return x, {"initial": perf[0], "final": perf[1], "nIterations": int(perf[2]),
                   "converged": bool(perf[3]), "singular": bool(perf[4]), "history": h}


def timed_solve(path, M, solver, precond, b, k1, k2, rel_tol, mg=None):
    """CPU timing of one preconditioner: (perf of the second solve, setup s,
    cold solve s (k1 iterations max), warm solve s (k2 iterations max))."""
    x = np.zeros(M.n)
    b = np.ascontiguousarray(b, np.float64)
    perf = np.zeros(6)
    sec = np.zeros(3)
    mga = None if mg is None else np.asarray(mg, np.int64)
    rc = path.fn("timed_solve")(M.h, SOLVER[solver], PRECOND[precond], _p(mga, _i64p), k1, k2, rel_tol,
                                _p(x), _p(b), _p(perf), _p(sec))
    if rc != 0:
        raise RuntimeError(f"timed_solve failed: {rc}")
    return ({"initial": perf[0], "final": perf[1], "nIterations": int(perf[2]), "converged": bool(perf[3])},
            float(sec[0]), float(sec[1]), float(sec[2]))


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run __graft_entry__.build()")
        _oracle = CpuSolvePath(ORACLE_LIB, "uo_")
    return _oracle


def reference():
    """The compiled reference, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        try:
            _ref = CpuSolvePath(REF_LIB, "ur_")
        except OSError:
            _ref = None
    return _ref
