"""ctypes access to oracle/_ref/libhonei_ref.so -- TEST INFRASTRUCTURE ONLY.

The library is HONEI's own tags::CPU code (compiled from the reference
sources by ``make -C oracle ref``).  It exists only where those sources were
present at build time; callers must skip when ``available()`` is False.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os

import numpy as np

from . import HERE, Ell

LIB_PATH = os.path.join(HERE, "_ref", "libhonei_ref.so")
_u64 = C.c_uint64
_p = C.c_void_p
_lib = None
_live: dict = {}


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB_PATH)
        for s in ("f", "d"):
            getattr(_lib, f"ref_ell_from_sm_{s}").restype = _p
            getattr(_lib, f"ref_q1_ell_{s}").restype = _p
        _lib.ref_ell_from_csr_d.restype = _p
        _lib.ref_prolmat_d.restype = _p
        _lib.ref_cg_solver_d.restype = _u64
        _lib.ref_pcg_jacobi_f.restype = _u64
        _lib.ref_pcg_jacobi_d.restype = _u64
        _lib.ref_mg_mixed.restype = _u64
        atexit.register(_free_all)
    return _lib


def _free_all():
    for h, s in list(_live.items()):
        getattr(_lib, f"ref_ell_free_{s}")(_p(h))
        _live.pop(h, None)


def _ptr(a):
    return a.ctypes.data_as(_p)


class RefEll:
    """Owning handle to a reference SparseMatrixELL<DT>."""

    def __init__(self, handle, sfx):
        self.h, self.s = handle, sfx
        _live[handle] = sfx

    def free(self):
        if self.h in _live:
            getattr(lib(), f"ref_ell_free_{self.s}")(_p(self.h))
            _live.pop(self.h, None)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def dtype(self):
        return np.float32 if self.s == "f" else np.float64

    def to_ell(self) -> Ell:
        v = [_u64() for _ in range(6)]
        getattr(lib(), f"ref_ell_info_{self.s}")(_p(self.h), *[C.byref(t) for t in v])
        rows, cols, stride, ncpr, T, size = (t.value for t in v)
        Aj = np.zeros(size, np.uint64)
        Ax = np.zeros(size, self.dtype)
        Arl = np.zeros(rows, np.uint64)
        getattr(lib(), f"ref_ell_copy_{self.s}")(_p(self.h), _ptr(Aj), _ptr(Ax), _ptr(Arl))
        return Ell(rows, cols, stride, ncpr, T, Aj, Ax, Arl)

    def product(self, x):
        x = np.ascontiguousarray(x, self.dtype)
        rows = self.to_ell().rows
        y = np.zeros(rows, self.dtype)
        getattr(lib(), f"ref_product_{self.s}")(_p(self.h), _ptr(x), _ptr(y))
        return y

    def defect(self, rhs, x):
        rhs = np.ascontiguousarray(rhs, self.dtype)
        x = np.ascontiguousarray(x, self.dtype)
        r = np.zeros(rhs.size, self.dtype)
        getattr(lib(), f"ref_defect_{self.s}")(_p(self.h), _ptr(rhs), _ptr(x), _ptr(r))
        return r


def ell_from_sm(rows, cols, rp, col, val, threads=1) -> RefEll:
    val = np.ascontiguousarray(val)
    s = "f" if val.dtype == np.float32 else "d"
    h = getattr(lib(), f"ref_ell_from_sm_{s}")(_u64(rows), _u64(cols), _ptr(np.ascontiguousarray(rp, np.uint64)),
                                                _ptr(np.ascontiguousarray(col, np.uint64)), _ptr(val), _u64(threads))
    return RefEll(h, s)


def ell_from_csr(rows, cols, rp, col, val, threads=1) -> RefEll:
    val = np.ascontiguousarray(val, np.float64)
    h = lib().ref_ell_from_csr_d(_u64(rows), _u64(cols), _ptr(np.ascontiguousarray(rp, np.uint64)),
                                 _ptr(np.ascontiguousarray(col, np.uint64)), _ptr(val), _u64(threads))
    return RefEll(h, "d")


def q1_ell(side, dtype=np.float64, bc=0, threads=1) -> RefEll:
    s = "f" if np.dtype(dtype) == np.float32 else "d"
    return RefEll(getattr(lib(), f"ref_q1_ell_{s}")(_u64(side), C.c_int(bc), _u64(threads)), s)


def prolmat(fine_side, threads=1) -> RefEll:
    return RefEll(lib().ref_prolmat_d(_u64(fine_side), _u64(threads)), "d")


def cg_solver(A: RefEll, b, x0, max_iters, eps):
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, np.float64, copy=True)
    used = lib().ref_cg_solver_d(_p(A.h), _ptr(b), _ptr(x), _u64(max_iters), C.c_double(eps))
    return x, used


def pcg_jacobi(A: RefEll, b, dinv, x0, max_iters, eps):
    """ConjugateGradients<tags::CPU, JAC>::value on the reference's ELL."""
    dt = A.dtype
    b = np.ascontiguousarray(b, dt)
    dinv = np.ascontiguousarray(dinv, dt)
    x = np.array(x0, dt, copy=True)
    sc = C.c_float if A.s == "f" else C.c_double
    used = getattr(lib(), f"ref_pcg_jacobi_{A.s}")(_p(A.h), _ptr(b), _ptr(x), _ptr(dinv), _u64(max_iters), sc(eps))
    return x, used


def cg_konv(A: RefEll, b, konv):
    b = np.ascontiguousarray(b, np.float32)
    x = np.zeros(b.size, np.float32)
    lib().ref_cg_konv_f(_p(A.h), _ptr(b), _ptr(x), C.c_double(konv))
    return x


def restriction(fine, mask=(2,) * 8):
    fine = np.ascontiguousarray(fine, np.float32)
    sf = int(round(np.sqrt(fine.size)))
    sc = (sf - 1) // 2 + 1
    c = np.zeros(sc * sc, np.float32)
    lib().ref_restriction_f(_ptr(c), _u64(c.size), _ptr(fine), _u64(fine.size), _ptr(np.asarray(mask, np.uint64)))
    return c


def prolongation(coarse, mask=(2,) * 8):
    coarse = np.ascontiguousarray(coarse, np.float32)
    sc = int(round(np.sqrt(coarse.size)))
    sf = 2 * (sc - 1) + 1
    f = np.zeros(sf * sf, np.float32)
    lib().ref_prolongation_f(_ptr(f), _u64(f.size), _ptr(coarse), _u64(coarse.size), _ptr(np.asarray(mask, np.uint64)))
    return f


def banded_q1(bands, x, rhs=None):
    """The reference's Product<tags::CPU>::value(BandedMatrixQx<Q1Type>, x), or
    Defect<tags::CPU> (rhs - A x) when rhs is given; bands[9][n] in the order
    LL, LD, LU, DL, DD, DU, UL, UD, UU."""
    bands = np.ascontiguousarray(bands)
    n = bands.shape[1]
    s = "f" if bands.dtype == np.float32 else "d"
    x = np.ascontiguousarray(x, bands.dtype)
    y = np.zeros(n, bands.dtype)
    r = None if rhs is None else np.ascontiguousarray(rhs, bands.dtype)
    getattr(lib(), f"ref_banded_q1_{s}")(_u64(n), _ptr(bands), None if r is None else _ptr(r), _ptr(x), _ptr(y))
    return y


def mg_mixed(max_level, b, x0, n_max_iter=4, n_pre=2, tol=1e-8, adapt=1.0, timed=False):
    """Multigrid<CPU, CPU, NONE, JAC, V, MIXED>::value of the reference on the
    Dirichlet Q1 problem; with timed=True also returns the seconds spent in
    the solve call alone (hierarchy set-up excluded)."""
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, np.float64, copy=True)
    inner = _u64()
    secs = C.c_double()
    outer = lib().ref_mg_mixed(_u64(max_level), _u64(n_max_iter), _u64(n_pre), C.c_double(tol),
                               C.c_double(adapt), _ptr(b), _ptr(x), C.byref(inner), C.byref(secs))
    if timed:
        return x, outer, inner.value, secs.value
    return x, outer, inner.value
