"""ctypes binding of the C-ABI in include/unap_hip.h (lib/libunap_hip.so).

Mirrors the reference's C++ plug-in API: LduMatrix (lduMatrix), Precond
(matrix::preconditioner), Smoother (matrix::smoother), Agglomeration
(lduAgglomeration) and Solver (matrix::solver with solve(x, A, b)).  Vectors
are numpy arrays (host) or objects exposing data_ptr() (device tensors).
There is no CPU fallback: every call runs the HIP library.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libunap_hip.so")

HOST, DEVICE = 0, 1
CYCLE_AUTO, CYCLE_REFERENCE, CYCLE_FIXED = -1, 0, 1
PRECOND = {"none": 0, "diag": 1, "DIC": 2, "DILU": 3, "AMG": 4}
SMOOTHER = {"GS": 1, "Cheby": 2}
SOLVER = {"PCG": 1, "PBiCGStab": 2, "MG": 3, "FCG": 4, "FBiCGStab": 5}

_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_intp = ctypes.POINTER(ctypes.c_int)


class AmgParams(ctypes.Structure):
    _fields_ = [("smoother", _i64), ("n_cells_coarsest", _i64), ("max_levels", _i64),
                ("n_vcycles", _i64), ("n_pre_sweeps", _i64), ("n_post_sweeps", _i64),
                ("n_finest_sweeps", _i64), ("precision", _i64), ("level_sweep_growth", _i64),
                ("pair_steps", _i64), ("cycle", _i64), ("scaling", _i64),
                ("cheby_eig_ratio", ctypes.c_double), ("cheby_boost", ctypes.c_double),
                ("cheby_eig_ratio_finest", ctypes.c_double)]


class Perf(ctypes.Structure):
    _fields_ = [("initial_residual", ctypes.c_double), ("final_residual", ctypes.c_double),
                ("previous_residual", ctypes.c_double), ("n_iterations", _i64),
                ("converged", ctypes.c_int32), ("singular", ctypes.c_int32)]

    def as_dict(self):
        return {"initial": self.initial_residual, "final": self.final_residual,
                "previous": self.previous_residual, "nIterations": int(self.n_iterations),
                "converged": bool(self.converged), "singular": bool(self.singular)}


# name -> (restype, argtypes); every function declared in include/unap_hip.h
SIGNATURES = {
    "unap_hip_last_error": (ctypes.c_char_p, []),
    "unap_hip_device_info": (ctypes.c_int, [ctypes.c_char_p, _i64, _i64p]),
    "unap_hip_set_reduction_order": (ctypes.c_int, [ctypes.c_int]),
    "unap_hip_set_graph_replay": (ctypes.c_int, [ctypes.c_int]),
    "unap_hip_ldu_create": (ctypes.c_int, [_i64, _i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(_vp)]),
    "unap_hip_ldu_destroy": (ctypes.c_int, [_vp]),
    "unap_hip_ldu_info": (ctypes.c_int, [_vp, _i64p, _i64p, _intp, _i64p, _i64p]),
    "unap_hip_ldu_set_coeffs": (ctypes.c_int, [_vp, _f64p, _f64p, _f64p]),
    "unap_hip_ldu_set_topology": (ctypes.c_int, [_vp, _i64, _i64p, _i64p]),
    "unap_hip_ldu_get": (ctypes.c_int, [_vp, _i64p, _i64p, _f64p, _f64p, _f64p]),
    "unap_hip_coo_to_ldu": (ctypes.c_int, [_i64, _i64, _i64p, _i64p, _f64p, ctypes.c_int, _i64, _i64p, _i64p, _i64p,
                                           _f64p, _f64p, _f64p]),
    "unap_hip_ldu_create_coo": (ctypes.c_int, [_i64, _i64, _i64p, _i64p, _f64p, ctypes.c_int, _i64,
                                               ctypes.POINTER(_vp)]),
    "unap_hip_ldu_create_csr": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, ctypes.c_int, _i64, ctypes.POINTER(_vp)]),
    "unap_hip_ldu_read_openfoam": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "unap_hip_ldu_write_openfoam": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "unap_hip_vector_read_openfoam": (ctypes.c_int, [ctypes.c_char_p, _i64, _f64p]),
    "unap_hip_vector_write_openfoam": (ctypes.c_int, [ctypes.c_char_p, _i64, _f64p]),
    "unap_hip_ldu_read_hypre": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "unap_hip_vector_read_hypre": (ctypes.c_int, [ctypes.c_char_p, _i64, _f64p]),
    "unap_hip_rcm_order": (ctypes.c_int, [_i64, _i64, _i64p, _i64p, _i64p]),
    "unap_hip_setup_self_check": (ctypes.c_int, [_vp, _f64p, _i64p]),
    "unap_hip_ldu_renumber": (ctypes.c_int, [_i64, _i64, _i64p, _i64p, _i64p, _f64p, _f64p, _f64p,
                                             _i64p, _i64p, _i64p, _f64p, _f64p, _f64p]),
    "unap_hip_ldu_spmv": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "unap_hip_ldu_spmv_noalloc": (ctypes.c_int, [_vp, _vp, _vp]),
    "unap_hip_amg_params_default": (ctypes.c_int, [ctypes.POINTER(AmgParams)]),
    "unap_hip_precond_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(AmgParams), _f64p, ctypes.POINTER(_vp)]),
    "unap_hip_precond_destroy": (ctypes.c_int, [_vp]),
    "unap_hip_precond_apply": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "unap_hip_precond_amg_levels": (ctypes.c_int, [_vp, _i64, _i64p, _i64p, _i64p]),
    "unap_hip_dense_inverse_spd": (ctypes.c_int, [_i64, _f64p, _f64p]),
    "unap_hip_precond_rd": (ctypes.c_int, [_vp, _f64p]),
    "unap_hip_smoother_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "unap_hip_smoother_destroy": (ctypes.c_int, [_vp]),
    "unap_hip_smoother_smooth": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, ctypes.c_int]),
    "unap_hip_smoother_init": (ctypes.c_int, [_vp]),
    "unap_hip_eigen_estimate": (ctypes.c_int, [_vp, _vp, _vp, _i64, ctypes.c_int, _f64p]),
    "unap_hip_aggl_create": (ctypes.c_int, [_vp, _f64p, _i64, _i64, ctypes.POINTER(_vp)]),
    "unap_hip_aggl_destroy": (ctypes.c_int, [_vp]),
    "unap_hip_aggl_nlevels": (ctypes.c_int, [_vp, _i64p]),
    "unap_hip_aggl_level_size": (ctypes.c_int, [_vp, _i64, _i64p, _i64p]),
    "unap_hip_aggl_restrict_addressing": (ctypes.c_int, [_vp, _i64, _i64p]),
    "unap_hip_aggl_face_restrict_addressing": (ctypes.c_int, [_vp, _i64, _i64p]),
    "unap_hip_aggl_coarse_matrix": (ctypes.c_int, [_vp, _i64, _i64p, _i64p, _f64p, _f64p, _f64p, _intp]),
    "unap_hip_aggl_restrict_field": (ctypes.c_int, [_vp, _i64, _vp, _vp, ctypes.c_int]),
    "unap_hip_aggl_prolong_field": (ctypes.c_int, [_vp, _i64, _vp, _vp, ctypes.c_int]),
    "unap_hip_solver_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "unap_hip_solver_destroy": (ctypes.c_int, [_vp]),
    "unap_hip_solver_set_precond": (ctypes.c_int, [_vp, _vp]),
    "unap_hip_solver_set_limits": (ctypes.c_int, [_vp, _i64, _i64, ctypes.c_double, ctypes.c_double]),
    "unap_hip_solver_set_mg": (ctypes.c_int, [_vp, ctypes.POINTER(AmgParams), _f64p]),
    "unap_hip_solver_solve": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.POINTER(Perf)]),
    "unap_hip_solver_history": (ctypes.c_int, [_vp, _f64p, _i64, _i64p]),
    "unap_hip_solver_last_ms": (ctypes.c_int, [_vp, _f64p]),
    "unap_hip_profile_kernel": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i64, _f64p, _f64p]),
    "unap_hip_batch_create": (ctypes.c_int, [_i64, _i64p, _i64p, _i64p, _i64p, _f64p, _f64p, _f64p, _i32p,
                                             ctypes.POINTER(_vp)]),
    "unap_hip_batch_destroy": (ctypes.c_int, [_vp]),
    "unap_hip_batch_solve": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _i64, ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(Perf)]),
    "unap_hip_batch_last_ms": (ctypes.c_int, [_vp, _f64p]),
    "unap_hip_batch_config": (ctypes.c_int, [_vp, _i64p, _i64p, _i64p]),
    "unap_hip_batch_set_config": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
}

_lib = None


class UnapHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"unap_hip error {code}: {msg}")
        self.code = code


def lib():
    """The loaded C-ABI library.  Raises when it is missing: there is no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise UnapHipError(rc, lib().unap_hip_last_error().decode())


def _p(a, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)


def _vec(v):
    """(pointer, where, keepalive) for a numpy array or a device tensor."""
    if hasattr(v, "data_ptr"):
        return ctypes.c_void_p(v.data_ptr()), DEVICE, v
    a = np.ascontiguousarray(v, dtype=np.float64)
    return a.ctypes.data_as(ctypes.c_void_p), HOST, a


REDUCE_TREE, REDUCE_REFERENCE = 0, 1


def set_reduction_order(order):
    """REDUCE_TREE (fast, default) or REDUCE_REFERENCE (the reference's serial sums)."""
    _check(lib().unap_hip_set_reduction_order(order))


def set_graph_replay(on):
    """Replay solver iterations from a HIP graph (True, default) or launch them kernel by kernel."""
    _check(lib().unap_hip_set_graph_replay(1 if on else 0))


def device_info():
    buf = ctypes.create_string_buffer(256)
    cus = _i64()
    _check(lib().unap_hip_device_info(buf, 256, ctypes.byref(cus)))
    return buf.value.decode(), int(cus.value)


_rt = None


def _runtime():
    """The HIP runtime the library itself is linked against (same process
    instance, resolved by soname after the library is loaded)."""
    global _rt
    if _rt is None:
        lib()
        R = ctypes.CDLL("libamdhip64.so.7")
        R.hipMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
        R.hipFree.argtypes = [_vp]
        R.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
        R.hipMemset.argtypes = [_vp, ctypes.c_int, ctypes.c_size_t]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize"):
            getattr(R, f).restype = ctypes.c_int
        _rt = R
    return _rt


class DeviceArray:
    """A float64 vector in device memory (hipMalloc), passed to the library
    as a DEVICE pointer.  Plumbing for callers without their own allocator."""

    def __init__(self, n, host=None):
        R = _runtime()
        self.n = int(n)
        self.ptr = _vp()
        if R.hipMalloc(ctypes.byref(self.ptr), max(8, 8 * self.n)) != 0:
            raise RuntimeError(f"hipMalloc of {8 * self.n} bytes failed")
        if host is None:
            self.zero()
        else:
            self.upload(host)

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, np.float64)
        return cls(a.size, host=a)

    def data_ptr(self):
        return self.ptr.value

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.n
        if _rt.hipMemcpy(self.ptr, a.ctypes.data_as(_vp), 8 * self.n, 1) != 0:
            raise RuntimeError("hipMemcpy host->device failed")

    def zero(self):
        if _rt.hipMemset(self.ptr, 0, 8 * self.n) != 0 or _rt.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")

    def numpy(self):
        out = np.empty(self.n)
        if _rt.hipMemcpy(out.ctypes.data_as(_vp), self.ptr, 8 * self.n, 2) != 0:
            raise RuntimeError("hipMemcpy device->host failed")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value and _rt is not None:
            _rt.hipFree(self.ptr)
            self.ptr = None


def synchronize():
    if _runtime().hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def coo_to_ldu(n, rows, cols, vals, symmetric=False, base=0):
    """(lowerAddr, upperAddr, lower, diag, upper) of the COO matrix, as coo2ldu gives them."""
    r = np.ascontiguousarray(rows, np.int64)
    c = np.ascontiguousarray(cols, np.int64)
    v = np.ascontiguousarray(vals, np.float64)
    nf = _i64()
    L = lib()
    _check(L.unap_hip_coo_to_ldu(n, r.size, _p(r, _i64p), _p(c, _i64p), _p(v), int(symmetric), base,
                                 ctypes.byref(nf), None, None, None, None, None))
    m = nf.value
    l = np.empty(m, np.int64)
    u = np.empty(m, np.int64)
    lo = np.empty(m)
    d = np.empty(n)
    up = np.empty(m)
    _check(L.unap_hip_coo_to_ldu(n, r.size, _p(r, _i64p), _p(c, _i64p), _p(v), int(symmetric), base,
                                 ctypes.byref(nf), _p(l, _i64p), _p(u, _i64p), _p(lo), _p(d), _p(up)))
    return l, u, lo, d, up


def read_vector_openfoam(path, n):
    out = np.empty(n)
    _check(lib().unap_hip_vector_read_openfoam(os.fsencode(path), n, _p(out)))
    return out


def write_vector_openfoam(path, v):
    v = np.ascontiguousarray(v, np.float64)
    _check(lib().unap_hip_vector_write_openfoam(os.fsencode(path), v.size, _p(v)))


def read_vector_hypre(path, n):
    out = np.empty(n)
    _check(lib().unap_hip_vector_read_hypre(os.fsencode(path), n, _p(out)))
    return out


def setup_self_check(A, weights=None):
    """Fields in which the device-built setup of A differs from the host-built
    one (0: identical)."""
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    m = _i64()
    _check(lib().unap_hip_setup_self_check(A.h, _p(w), ctypes.byref(m)))
    return m.value


def rcm_order(n, lower_addr, upper_addr):
    """Reverse Cuthill-McKee order (cooRCM.cpp): the new label of every cell."""
    l = np.ascontiguousarray(lower_addr, np.int64)
    u = np.ascontiguousarray(upper_addr, np.int64)
    out = np.empty(n, np.int64)
    _check(lib().unap_hip_rcm_order(n, l.size, _p(l, _i64p), _p(u, _i64p), _p(out, _i64p)))
    return out


def renumber(n, lower_addr, upper_addr, new_of_old, diag=None, upper=None, lower=None):
    """rcmLDU_rewrite for a given cell order: (lowerAddr, upperAddr,
    new_face_of_old, lower, diag, upper) of the renumbered matrix (faces
    turned to owner < neighbour and sorted; ~k marks a face that turned;
    lower None for a symmetric matrix)."""
    l = np.ascontiguousarray(lower_addr, np.int64)
    u = np.ascontiguousarray(upper_addr, np.int64)
    p = np.ascontiguousarray(new_of_old, np.int64)
    nf = l.size
    d = None if diag is None else np.ascontiguousarray(diag, np.float64)
    up = None if upper is None else np.ascontiguousarray(upper, np.float64)
    lo = None if lower is None else np.ascontiguousarray(lower, np.float64)
    lo_out = np.empty(nf) if (up is not None and lo is not None) else None
    up_out = np.empty(nf) if up is not None else None
    d_out = np.empty(n) if d is not None else None
    l_out, u_out, fmap = np.empty(nf, np.int64), np.empty(nf, np.int64), np.empty(nf, np.int64)
    _check(lib().unap_hip_ldu_renumber(n, nf, _p(l, _i64p), _p(u, _i64p), _p(p, _i64p), _p(lo), _p(d), _p(up),
                                       _p(l_out, _i64p), _p(u_out, _i64p), _p(fmap, _i64p), _p(lo_out), _p(d_out),
                                       _p(up_out)))
    return l_out, u_out, fmap, lo_out, d_out, up_out


def dense_inverse_spd(a):
    """Inverse of a dense SPD matrix by the GPU's blocked Gauss-Jordan."""
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    out = np.empty((n, n))
    _check(lib().unap_hip_dense_inverse_spd(n, _p(a), _p(out)))
    return out


def amg_params(**kw):
    p = AmgParams()
    _check(lib().unap_hip_amg_params_default(ctypes.byref(p)))
    for k, v in kw.items():
        setattr(p, k, float(v) if k in ("cheby_eig_ratio", "cheby_boost", "cheby_eig_ratio_finest") else int(v))
    return p


class LduMatrix:
    """lduMatrix(nCells, lowerAddr, upperAddr, lower, diag, upper) on the GPU."""

    def __init__(self, n, lower_addr, upper_addr, lower, diag, upper):
        L = lib()
        self._l = np.ascontiguousarray(lower_addr, np.int64)
        self._u = np.ascontiguousarray(upper_addr, np.int64)
        self._up = np.ascontiguousarray(upper, np.float64)
        self._lo = None if lower is None or lower is upper else np.ascontiguousarray(lower, np.float64)
        self._d = np.ascontiguousarray(diag, np.float64)
        self.n = int(n)
        h = _vp()
        _check(L.unap_hip_ldu_create(self.n, self._u.shape[0], _p(self._l, _i64p), _p(self._u, _i64p),
                                     _p(self._lo) if self._lo is not None else _p(self._up),
                                     _p(self._d), _p(self._up), ctypes.byref(h)))
        self.h = h

    @classmethod
    def from_system(cls, s):
        return cls(s.n, s.l, s.u, None if s.symmetric else s.lower, s.diag, s.upper)

    @classmethod
    def _from_handle(cls, h):
        A = cls.__new__(cls)
        A.h = h
        n, nf, sy = _i64(), _i64(), ctypes.c_int()
        _check(lib().unap_hip_ldu_info(h, ctypes.byref(n), ctypes.byref(nf), ctypes.byref(sy), None, None))
        A.n = n.value
        A._l, A._u, A._lo, A._d, A._up = A._arrays(nf.value, bool(sy.value))
        return A

    def _arrays(self, nf, symm):
        l = np.empty(nf, np.int64)
        u = np.empty(nf, np.int64)
        lo = np.empty(nf)
        d = np.empty(self.n)
        up = np.empty(nf)
        _check(lib().unap_hip_ldu_get(self.h, _p(l, _i64p), _p(u, _i64p), _p(lo), _p(d), _p(up)))
        return l, u, (None if symm else lo), d, up

    def arrays(self):
        """(lowerAddr, upperAddr, lower, diag, upper) host copies; lower is None when symmetric."""
        info = self.info(levels=False)
        return self._arrays(info["nfaces"], info["symmetric"])

    @classmethod
    def from_coo(cls, n, rows, cols, vals, symmetric=False, base=0):
        """coo2ldu after reorderCOO (coo2ldu.cpp, matrixConversion.cpp)."""
        r = np.ascontiguousarray(rows, np.int64)
        c = np.ascontiguousarray(cols, np.int64)
        v = np.ascontiguousarray(vals, np.float64)
        h = _vp()
        _check(lib().unap_hip_ldu_create_coo(n, r.size, _p(r, _i64p), _p(c, _i64p), _p(v), int(symmetric), base,
                                             ctypes.byref(h)))
        return cls._from_handle(h)

    @classmethod
    def from_csr(cls, n, row_start, cols, vals, symmetric=False, base=0):
        """csr2ldu (coo2ldu.cpp:117-148)."""
        rs = np.ascontiguousarray(row_start, np.int64)
        c = np.ascontiguousarray(cols, np.int64)
        v = np.ascontiguousarray(vals, np.float64)
        h = _vp()
        _check(lib().unap_hip_ldu_create_csr(n, _p(rs, _i64p), _p(c, _i64p), _p(v), int(symmetric), base,
                                             ctypes.byref(h)))
        return cls._from_handle(h)

    @classmethod
    def read_openfoam(cls, path):
        h = _vp()
        _check(lib().unap_hip_ldu_read_openfoam(os.fsencode(path), ctypes.byref(h)))
        return cls._from_handle(h)

    @classmethod
    def read_hypre(cls, path):
        h = _vp()
        _check(lib().unap_hip_ldu_read_hypre(os.fsencode(path), ctypes.byref(h)))
        return cls._from_handle(h)

    def write_openfoam(self, path):
        _check(lib().unap_hip_ldu_write_openfoam(self.h, os.fsencode(path)))

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.unap_hip_ldu_destroy(self.h)
            self.h = None

    def info(self, levels=True):
        """Sizes and symmetry; with levels, the forward / backward sweep level
        counts too (computed on the first request: a pass over every row)."""
        n, nf, fl, bl = _i64(), _i64(), _i64(), _i64()
        sy = ctypes.c_int()
        lv = (ctypes.byref(fl), ctypes.byref(bl)) if levels else (None, None)
        _check(lib().unap_hip_ldu_info(self.h, ctypes.byref(n), ctypes.byref(nf), ctypes.byref(sy), *lv))
        out = {"n": n.value, "nfaces": nf.value, "symmetric": bool(sy.value)}
        if levels:
            out.update(fwd_levels=fl.value, bwd_levels=bl.value)
        return out

    def set_coeffs(self, diag, upper, lower=None):
        """setMatrixCoeffients: new values, same addressing (lower None: symmetric)."""
        self._d = np.ascontiguousarray(diag, np.float64)
        self._up = np.ascontiguousarray(upper, np.float64)
        self._lo = None if lower is None or lower is upper else np.ascontiguousarray(lower, np.float64)
        _check(lib().unap_hip_ldu_set_coeffs(self.h, _p(self._d), _p(self._up), _p(self._lo)))

    def set_topology(self, lower_addr, upper_addr):
        """setMatrixTopology: new addressing (same cells); coefficients untouched."""
        self._l = np.ascontiguousarray(lower_addr, np.int64)
        self._u = np.ascontiguousarray(upper_addr, np.int64)
        _check(lib().unap_hip_ldu_set_topology(self.h, self._u.shape[0], _p(self._l, _i64p), _p(self._u, _i64p)))

    def spmv(self, x, out=None):
        xp, where, keep = _vec(x)
        if where == HOST:
            y = np.empty(self.n)
            _check(lib().unap_hip_ldu_spmv(self.h, xp, y.ctypes.data_as(ctypes.c_void_p), HOST))
            return y
        _check(lib().unap_hip_ldu_spmv(self.h, xp, ctypes.c_void_p(out.data_ptr()), DEVICE))
        return out


    def spmv_noalloc(self, x, out):
        """A x into out, both device arrays, with no device allocation."""
        _check(lib().unap_hip_ldu_spmv_noalloc(self.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr())))
        return out


KERNELS = {"amul": 1, "cheby": 2, "gs": 3, "dic": 4, "residual": 5, "axpy_norm": 6, "dot": 7, "copy": 8}


def profile_kernel(A, kernel, precision=64, reps=20):
    """(mean ms per launch, algorithmic bytes per launch) of one kernel on A."""
    ms, by = ctypes.c_double(), ctypes.c_double()
    _check(lib().unap_hip_profile_kernel(A.h, KERNELS[kernel], precision, reps, ctypes.byref(ms), ctypes.byref(by)))
    return ms.value, by.value


class Precond:
    def __init__(self, A, kind, amg=None, weights=None, precision=None):
        """precision=32 with DIC / DILU: FP32 sweeps (for FCG and PBiCGStab)."""
        self.A = A
        if precision is not None and amg is None:
            amg = amg_params(precision=precision)
        h = _vp()
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        _check(lib().unap_hip_precond_create(A.h, PRECOND[kind], None if amg is None else ctypes.byref(amg),
                                             _p(w), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.unap_hip_precond_destroy(self.h)
            self.h = None

    def apply(self, r):
        rp, where, keep = _vec(r)
        w = np.empty(self.A.n)
        _check(lib().unap_hip_precond_apply(self.h, rp, w.ctypes.data_as(ctypes.c_void_p), HOST))
        return w

    def rD(self):
        out = np.empty(self.A.n)
        _check(lib().unap_hip_precond_rd(self.h, _p(out)))
        return out

    def amg_levels(self):
        """[(n_cells, n_faces)] of the AMG hierarchy, level 0 = the matrix."""
        L = _i64()
        _check(lib().unap_hip_precond_amg_levels(self.h, 0, ctypes.byref(L), None, None))
        n = np.zeros(L.value, np.int64)
        f = np.zeros(L.value, np.int64)
        _check(lib().unap_hip_precond_amg_levels(self.h, L.value, ctypes.byref(L), _p(n, _i64p), _p(f, _i64p)))
        return [(int(a), int(b)) for a, b in zip(n, f)]


class Smoother:
    def __init__(self, kind):
        h = _vp()
        _check(lib().unap_hip_smoother_create(SMOOTHER[kind], ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.unap_hip_smoother_destroy(self.h)
            self.h = None

    def smooth(self, A, x, b, nsweeps):
        x = np.array(x, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        _check(lib().unap_hip_smoother_smooth(self.h, A.h, x.ctypes.data_as(ctypes.c_void_p),
                                              b.ctypes.data_as(ctypes.c_void_p), nsweeps, HOST))
        return x

    def init(self):
        _check(lib().unap_hip_smoother_init(self.h))


def eigen_estimate(A, x, b, npcg=10):
    x = np.array(x, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    lam = ctypes.c_double()
    _check(lib().unap_hip_eigen_estimate(A.h, x.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
                                         npcg, HOST, ctypes.byref(lam)))
    return lam.value, x


class Agglomeration:
    def __init__(self, A, weights, max_levels=50, n_coarsest=50):
        self.A = A
        w = np.ascontiguousarray(weights, np.float64)
        h = _vp()
        _check(lib().unap_hip_aggl_create(A.h, _p(w), max_levels, n_coarsest, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.unap_hip_aggl_destroy(self.h)
            self.h = None

    def nlevels(self):
        n = _i64()
        _check(lib().unap_hip_aggl_nlevels(self.h, ctypes.byref(n)))
        return n.value

    def level(self, lev, fine_n, fine_nf, values=True):
        """Maps and coarse addressing of one level; with values, its Galerkin
        coefficients too (formed on the GPU)."""
        nc, nf = _i64(), _i64()
        L = lib()
        _check(L.unap_hip_aggl_level_size(self.h, lev, ctypes.byref(nc), ctypes.byref(nf)))
        nc, nf = nc.value, nf.value
        ra = np.empty(fine_n, np.int64)
        fa = np.empty(fine_nf, np.int64)
        _check(L.unap_hip_aggl_restrict_addressing(self.h, lev, _p(ra, _i64p)))
        _check(L.unap_hip_aggl_face_restrict_addressing(self.h, lev, _p(fa, _i64p)))
        l = np.empty(nf, np.int64)
        u = np.empty(nf, np.int64)
        lo = np.empty(nf) if values else None
        up = np.empty(nf) if values else None
        d = np.empty(nc) if values else None
        sy = ctypes.c_int()
        _check(L.unap_hip_aggl_coarse_matrix(self.h, lev, _p(l, _i64p), _p(u, _i64p), _p(lo), _p(d), _p(up),
                                             ctypes.byref(sy)))
        out = {"n": nc, "nfaces": nf, "restrict": ra, "face_restrict": fa, "l": l, "u": u,
               "symmetric": bool(sy.value)}
        if values:
            out.update(lower=lo, diag=d, upper=up)
        return out

    def levels(self, values=True):
        out = []
        n, nf = self.A.n, self.A._u.shape[0]
        for lev in range(self.nlevels()):
            lv = self.level(lev, n, nf, values)
            out.append(lv)
            n, nf = lv["n"], lv["nfaces"]
        return out

    def restrict_field(self, lev, ff, nc):
        ff = np.ascontiguousarray(ff, np.float64)
        cf = np.empty(nc)
        _check(lib().unap_hip_aggl_restrict_field(self.h, lev, ff.ctypes.data_as(ctypes.c_void_p),
                                                  cf.ctypes.data_as(ctypes.c_void_p), HOST))
        return cf

    def prolong_field(self, lev, cf, nf):
        cf = np.ascontiguousarray(cf, np.float64)
        ff = np.empty(nf)
        _check(lib().unap_hip_aggl_prolong_field(self.h, lev, cf.ctypes.data_as(ctypes.c_void_p),
                                                 ff.ctypes.data_as(ctypes.c_void_p), HOST))
        return ff


class Solver:
    """matrix::solver: PCG / PBiCGStab / MG / FCG / FBiCGStab with solve(x, A, b)."""

    def __init__(self, kind, precond=None, max_iter=1000, min_iter=0, rel_tol=1e-6, tol=0.0, mg=None,
                 mg_weights=None):
        L = lib()
        h = _vp()
        _check(L.unap_hip_solver_create(SOLVER[kind], ctypes.byref(h)))
        self.h = h
        self.precond = precond
        if precond is not None:
            _check(L.unap_hip_solver_set_precond(h, precond.h))
        _check(L.unap_hip_solver_set_limits(h, max_iter, min_iter, rel_tol, tol))
        self._mgw = None if mg_weights is None else np.ascontiguousarray(mg_weights, np.float64)
        if mg is not None:
            _check(L.unap_hip_solver_set_mg(h, ctypes.byref(mg), _p(self._mgw)))

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.unap_hip_solver_destroy(self.h)
            self.h = None

    def solve(self, A, b, x0=None, x=None):
        """Host arrays: returns (x, perf).  Device tensors: pass x (in/out) and b."""
        perf = Perf()
        if hasattr(b, "data_ptr"):
            _check(lib().unap_hip_solver_solve(self.h, A.h, ctypes.c_void_p(x.data_ptr()),
                                               ctypes.c_void_p(b.data_ptr()), DEVICE, ctypes.byref(perf)))
            return x, perf.as_dict()
        xx = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
        bb = np.ascontiguousarray(b, np.float64)
        _check(lib().unap_hip_solver_solve(self.h, A.h, xx.ctypes.data_as(ctypes.c_void_p),
                                           bb.ctypes.data_as(ctypes.c_void_p), HOST, ctypes.byref(perf)))
        d = perf.as_dict()
        d["history"] = self.history()
        return xx, d

    def history(self):
        n = _i64()
        _check(lib().unap_hip_solver_history(self.h, None, 0, ctypes.byref(n)))
        out = np.empty(n.value)
        _check(lib().unap_hip_solver_history(self.h, _p(out), n.value, ctypes.byref(n)))
        return out

    def last_ms(self):
        v = ctypes.c_double()
        _check(lib().unap_hip_solver_last_ms(self.h, ctypes.byref(v)))
        return v.value


class Batch:
    """Many independent LDU systems solved in one persistent kernel launch."""

    def __init__(self, systems, solvers):
        ns = len(systems)
        cell_off = np.zeros(ns + 1, np.int64)
        face_off = np.zeros(ns + 1, np.int64)
        for i, s in enumerate(systems):
            cell_off[i + 1] = cell_off[i] + s.n
            face_off[i + 1] = face_off[i] + s.nfaces
        self.l = np.concatenate([s.l for s in systems]).astype(np.int64)
        self.u = np.concatenate([s.u for s in systems]).astype(np.int64)
        self.up = np.concatenate([s.upper for s in systems])
        self.lo = np.concatenate([s.lower_or_upper() for s in systems])
        self.d = np.concatenate([s.diag for s in systems])
        self.cell_off, self.face_off = cell_off, face_off
        self.kinds = np.array([SOLVER[k] for k in solvers], np.int32)
        h = _vp()
        _check(lib().unap_hip_batch_create(ns, _p(cell_off, _i64p), _p(face_off, _i64p), _p(self.l, _i64p),
                                           _p(self.u, _i64p), _p(self.lo), _p(self.d), _p(self.up),
                                           _p(self.kinds, _i32p), ctypes.byref(h)))
        self.h = h
        self.ns = ns

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.unap_hip_batch_destroy(self.h)
            self.h = None

    def solve(self, b, x=None, max_iter=1000, rel_tol=1e-8, tol=0.0):
        perfs = (Perf * self.ns)()
        if hasattr(b, "data_ptr"):
            _check(lib().unap_hip_batch_solve(self.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                              DEVICE, max_iter, rel_tol, tol, perfs))
            return x, [p.as_dict() for p in perfs]
        xx = np.zeros(int(self.cell_off[-1])) if x is None else np.array(x, np.float64)
        bb = np.ascontiguousarray(b, np.float64)
        _check(lib().unap_hip_batch_solve(self.h, xx.ctypes.data_as(ctypes.c_void_p),
                                          bb.ctypes.data_as(ctypes.c_void_p), HOST, max_iter, rel_tol, tol, perfs))
        return xx, [p.as_dict() for p in perfs]

    def last_ms(self):
        v = ctypes.c_double()
        _check(lib().unap_hip_batch_last_ms(self.h, ctypes.byref(v)))
        return v.value

    def config(self):
        """(threads per workgroup, workgroups per CU, rows per wavefront pass)."""
        bs, wg, rw = _i64(), _i64(), _i64()
        _check(lib().unap_hip_batch_config(self.h, ctypes.byref(bs), ctypes.byref(wg), ctypes.byref(rw)))
        return bs.value, wg.value, rw.value

    def set_config(self, block_size, wg_per_cu, rows_per_wavefront=256):
        """Fix the launch shape (the autotuner then does not run)."""
        _check(lib().unap_hip_batch_set_config(self.h, block_size, wg_per_cu, rows_per_wavefront))
