"""Seeded synthetic LDU systems (Python side of include/unap_meshgen.h).

The heavy lifting (permutation, face sort, coefficients) runs in
lib/libunap_meshgen.so; this module wraps it with numpy arrays and defines the
five workloads the solve path is measured on:

  C1  32^3 constant 7-point Poisson, PCG + DIC
  C2  128^3 log-uniform (1e-3..1) Poisson, PCG + agglomeration AMG (GS)
  C3  128^3 convection-diffusion (Peclet <= 5), PBiCGStab + DILU
  C4  256^3 log-uniform (1e-4..1) Poisson, AMG-PCG, mixed precision vs FP64
  C5  512 uneven box systems (8^3..48^3), 60% SPD / 40% convection-diffusion

Seeds: base seed 1000 + c for workload c; the generator derives its
sub-streams (permutation, coefficients, right-hand side, fluxes) from it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libunap_meshgen.so")
_lib = None

POISSON_CONST = 1
POISSON_LOGUNIF = 2
CONVDIFF = 3

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(_LIB_PATH)
        lib.unap_meshgen_box.restype = ctypes.c_int
        lib.unap_meshgen_box.argtypes = [ctypes.c_int64] * 3 + [
            ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
            _i64p, _i64p, _f64p, _f64p, _f64p, _f64p, _i64p]
        lib.unap_meshgen_nfaces.restype = ctypes.c_int64
        lib.unap_meshgen_nfaces.argtypes = [ctypes.c_int64] * 3
        lib.unap_meshgen_splitmix64.restype = ctypes.c_uint64
        lib.unap_meshgen_splitmix64.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        _lib = lib
    return _lib


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


class LduSystem:
    """One LDU system: n cells, f faces, owner/neighbour addressing, coefficients, rhs."""

    def __init__(self, n, l, u, lower, diag, upper, b, symmetric, dims=None, perm=None):
        self.n = int(n)
        self.l, self.u = l, u
        self.lower, self.diag, self.upper = lower, diag, upper
        self.b = b
        self.symmetric = bool(symmetric)
        self.dims = dims
        self.perm = perm

    @property
    def nfaces(self):
        return int(self.u.shape[0])

    def lower_or_upper(self):
        return self.upper if self.symmetric else self.lower

    def dense(self):
        """Dense matrix (small systems only; used by tests)."""
        A = np.zeros((self.n, self.n))
        A[np.arange(self.n), np.arange(self.n)] = self.diag
        A[self.l, self.u] = self.upper
        A[self.u, self.l] = self.lower_or_upper()
        return A


def box_system(nx, ny, nz, kind, p0=0.0, p1=0.0, shift=0.0, seed=0, with_perm=False):
    lib = _load()
    n = nx * ny * nz
    nf = int(lib.unap_meshgen_nfaces(nx, ny, nz))
    l = np.empty(nf, np.int64)
    u = np.empty(nf, np.int64)
    upper = np.empty(nf, np.float64)
    lower = np.empty(nf, np.float64) if kind == CONVDIFF else None
    diag = np.empty(n, np.float64)
    b = np.empty(n, np.float64)
    perm = np.empty(n, np.int64) if with_perm else None
    rc = lib.unap_meshgen_box(nx, ny, nz, kind, p0, p1, shift, seed,
                              _ptr(l, _i64p), _ptr(u, _i64p), _ptr(lower, _f64p),
                              _ptr(diag, _f64p), _ptr(upper, _f64p), _ptr(b, _f64p),
                              _ptr(perm, _i64p))
    if rc != 0:
        raise ValueError(f"unap_meshgen_box failed with {rc}")
    return LduSystem(n, l, u, lower, diag, upper, b, kind != CONVDIFF, (nx, ny, nz), perm)


def config_system(c, N=None, seed=None):
    """The box system of workload C1..C4, optionally at another edge length N."""
    s = 1000 + c if seed is None else seed
    if c == 1:
        N = 32 if N is None else N
        return box_system(N, N, N, POISSON_CONST, 0.0, 0.0, 1e-6, s)
    if c == 2:
        N = 128 if N is None else N
        return box_system(N, N, N, POISSON_LOGUNIF, -3.0, 0.0, 1e-6, s)
    if c == 3:
        N = 128 if N is None else N
        return box_system(N, N, N, CONVDIFF, 5.0, 0.0, 1e-3, s)
    if c == 4:
        N = 256 if N is None else N
        return box_system(N, N, N, POISSON_LOGUNIF, -4.0, 0.0, 1e-6, s)
    raise ValueError(c)


class SplitMix64:
    """Python twin of the generator's SplitMix64 stream (for small draws)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53

    def below(self, m):
        return (self.next() * m) >> 64


def batch_specs(nsys=512, lo=8, hi=48, spd_fraction=0.6, seed=1005):
    """Workload C5: per system (nx, ny, nz) each uniform in {lo..hi}; after a
    seeded Fisher-Yates shuffle the first spd_fraction are SPD log-uniform
    Poisson (PCG + DIC), the rest convection-diffusion (PBiCGStab + DILU).
    Returns a list of (nx, ny, nz, kind, sub_seed)."""
    g = SplitMix64(seed)
    dims = [tuple(lo + g.below(hi - lo + 1) for _ in range(3)) for _ in range(nsys)]
    order = list(range(nsys))
    for i in range(nsys - 1, 0, -1):
        j = g.below(i + 1)
        order[i], order[j] = order[j], order[i]
    nspd = int(round(spd_fraction * nsys))
    kinds = [None] * nsys
    for rank, s in enumerate(order):
        kinds[s] = POISSON_LOGUNIF if rank < nspd else CONVDIFF
    return [(d[0], d[1], d[2], kinds[s], seed * 1000003 + s) for s, d in enumerate(dims)]


def batch_system(spec):
    nx, ny, nz, kind, s = spec
    if kind == POISSON_LOGUNIF:
        return box_system(nx, ny, nz, POISSON_LOGUNIF, -3.0, 0.0, 1e-6, s)
    return box_system(nx, ny, nz, CONVDIFF, 5.0, 0.0, 1e-3, s)
