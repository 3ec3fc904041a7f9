"""CPU oracle for HONEI's ELL sparse-solve path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``honei_amd``) never imports, links or calls anything here.

The numerics live in ``honei_oracle.cpp`` (a clean-room restatement of the
reference's ``tags::CPU`` operators, each function citing the reference
file:line it follows); this module loads it with ctypes and adds the plain
file formats of the reference (``math/matrix_io.hh:433-584``,
``math/vector_io.hh:52-250``) in numpy.

Pinning (see tests/test_oracle_*.py): CSR->ELL against the reference's own
``5pt_10x10`` pair and the ``poisson/*.ell`` hierarchy (bit-exact), products
against dense brute force, transfers against the 25x9 prolongation table and
the geometric == PROLMAT invariant, solvers against the FEAT solution
``poisson/sol_6`` and against closed forms, and -- where the reference builds
in this container -- against ``oracle/_ref`` (the reference's own code).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_u64 = C.c_uint64
_p = C.c_void_p
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    for s, ct in (("f", C.c_float), ("d", C.c_double)):
        getattr(L, f"oracle_dot_{s}").restype = ct
        getattr(L, f"oracle_norm2_{s}").restype = ct
        for fn in ("oracle_scaled_sum2_", "oracle_scaled_sum3_", "oracle_scale_"):
            getattr(L, fn + s).argtypes = None
        getattr(L, f"oracle_banded_q1_to_csr_{s}").restype = _u64
        getattr(L, f"oracle_cg_konv_{s}").restype = _u64
    L.oracle_prolmat_csr.restype = _u64
    L.oracle_mg_fixed_d.restype = _u64
    L.oracle_mg_fixed_f.restype = _u64
    L.oracle_mg_mixed.restype = _u64
    L.oracle_mg_oplist_d.restype = _u64


def _ptr(a):
    return a.ctypes.data_as(_p)


def _sfx(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f"
    if dt == np.float64:
        return "d"
    raise TypeError(f"oracle supports float32/float64, not {dt}")


def _scalar(dtype, v):
    return C.c_float(v) if np.dtype(dtype) == np.float32 else C.c_double(v)


# ---------------------------------------------------------------------------
# ELL container (la/sparse_matrix_ell.hh:44-129)
# ---------------------------------------------------------------------------
@dataclass
class Ell:
    rows: int
    cols: int
    stride: int
    ncpr: int
    threads: int
    Aj: np.ndarray      # uint64 [ncpr*stride]
    Ax: np.ndarray      # DT     [ncpr*stride]
    Arl: np.ndarray     # uint64 [rows]

    @property
    def dtype(self):
        return self.Ax.dtype

    def astype(self, dtype) -> "Ell":
        return Ell(self.rows, self.cols, self.stride, self.ncpr, self.threads,
                   self.Aj, self.Ax.astype(dtype), self.Arl)

    def used_elements(self) -> int:
        return int(np.count_nonzero(self.Ax))

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.rows, self.cols), dtype=np.float64)
        T = self.threads
        for r in range(self.rows):
            for k in range(int(self.Arl[r])):
                for t in range(T):
                    s = r * T + t + k * self.stride
                    if self.Ax[s] != 0:
                        d[r, int(self.Aj[s])] = self.Ax[s]
        return d

    def to_csr(self):
        """SparseMatrix(ELL) (la/sparse_matrix.hh:157-185): non-zeros, sorted."""
        T, S = self.threads, self.stride
        rp = [0]
        cols, vals = [], []
        nch = len(self.Ax) // S if S else 0
        for r in range(self.rows):
            entries = {}
            for k in range(nch):
                for t in range(T):
                    s = r * T + t + k * S
                    if self.Ax[s] != 0:
                        entries[int(self.Aj[s])] = self.Ax[s]
            for c in sorted(entries):
                cols.append(c)
                vals.append(entries[c])
            rp.append(len(cols))
        return (np.asarray(rp, np.uint64), np.asarray(cols, np.uint64),
                np.asarray(vals, dtype=self.Ax.dtype))

    def c_struct(self):
        return OracleEll(self.rows, self.cols, self.stride, self.ncpr, self.threads,
                         _ptr(self.Aj), _ptr(self.Ax), _ptr(self.Arl))


def ell_from_sm(rows, cols, row_ptr, col, val, threads=1) -> Ell:
    """SparseMatrixELL(const SparseMatrix&) (la/sparse_matrix_ell-impl.hh:75-122)."""
    row_ptr = np.ascontiguousarray(row_ptr, np.uint64)
    col = np.ascontiguousarray(col, np.uint64)
    val = np.ascontiguousarray(val)
    stride, ncpr = _u64(), _u64()
    lib().oracle_ell_shape_sm(_u64(rows), _ptr(row_ptr), _u64(threads), C.byref(stride), C.byref(ncpr))
    size = stride.value * ncpr.value
    Aj = np.zeros(size, np.uint64)
    Ax = np.zeros(size, val.dtype)
    Arl = np.zeros(rows, np.uint64)
    getattr(lib(), f"oracle_ell_fill_sm_{_sfx(val.dtype)}")(
        _u64(rows), _ptr(row_ptr), _ptr(col), _ptr(val), _u64(threads),
        _u64(stride.value), _u64(ncpr.value), _ptr(Aj), _ptr(Ax), _ptr(Arl))
    return Ell(rows, cols, stride.value, ncpr.value, threads, Aj, Ax, Arl)


def ell_from_csr(rows, cols, Ar, Aj_csr, Ax_csr, threads=1, blocksize=1) -> Ell:
    """SparseMatrixELL(const SparseMatrixCSR&) (la/sparse_matrix_ell-impl.hh:124-183)."""
    Ar = np.ascontiguousarray(Ar, np.uint64)
    Aj_csr = np.ascontiguousarray(Aj_csr, np.uint64)
    Ax_csr = np.ascontiguousarray(Ax_csr)
    s = _sfx(Ax_csr.dtype)
    stride, ncpr = _u64(), _u64()
    getattr(lib(), f"oracle_ell_shape_csr_{s}")(
        _u64(rows), _ptr(Ar), _ptr(Ax_csr), _u64(blocksize), _u64(threads),
        C.byref(stride), C.byref(ncpr))
    size = stride.value * ncpr.value
    Aj = np.zeros(size, np.uint64)
    Ax = np.zeros(size, Ax_csr.dtype)
    Arl = np.zeros(rows, np.uint64)
    getattr(lib(), f"oracle_ell_fill_csr_{s}")(
        _u64(rows), _ptr(Ar), _ptr(Aj_csr), _ptr(Ax_csr), _u64(blocksize), _u64(threads),
        _u64(stride.value), _u64(ncpr.value), _ptr(Aj), _ptr(Ax), _ptr(Arl))
    return Ell(rows, cols, stride.value, ncpr.value, threads, Aj, Ax, Arl)


def ell_from_arrays(rows, cols, stride, ncpr, Aj, Ax, threads=1) -> Ell:
    """SparseMatrixELL(rows, columns, stride, ncpr, Aj, Ax, threads): row
    lengths derived from Ax (la/sparse_matrix_ell-impl.hh:61-73, 186-208)."""
    Aj = np.ascontiguousarray(Aj, np.uint64)
    Ax = np.ascontiguousarray(Ax)
    Arl = np.zeros(rows, np.uint64)
    getattr(lib(), f"oracle_ell_row_length_{_sfx(Ax.dtype)}")(
        _u64(rows), _u64(stride), _u64(threads), _u64(Ax.size), _ptr(Ax), _ptr(Arl))
    return Ell(rows, cols, stride, ncpr, threads, Aj, Ax, Arl)


# ---------------------------------------------------------------------------
# Operators
# ---------------------------------------------------------------------------
def _ell_args(A: Ell):
    return (_u64(A.rows), _u64(A.stride), _u64(A.threads), _ptr(A.Aj), _ptr(A.Ax), _ptr(A.Arl))


def product(A: Ell, x, row_start=0, row_end=0):
    """Product<tags::CPU>::value(y, ELL, x) (la/product.hh:148-206)."""
    x = np.ascontiguousarray(x, A.dtype)
    y = np.zeros(A.rows, A.dtype)
    getattr(lib(), f"oracle_product_ell_{_sfx(A.dtype)}")(
        _ptr(y), _ptr(x), *_ell_args(A), _u64(row_start), _u64(row_end))
    return y


def product_mc(A: Ell, x, parts):
    """mc::Product row-block split (la/product.hh:2114-2150)."""
    x = np.ascontiguousarray(x, A.dtype)
    y = np.zeros(A.rows, A.dtype)
    getattr(lib(), f"oracle_product_ell_mc_{_sfx(A.dtype)}")(
        _ptr(y), _ptr(x), *_ell_args(A), C.c_int(parts))
    return y


def defect(A: Ell, rhs, x):
    """Defect<tags::CPU>::value(r, rhs, ELL, x) (math/defect.hh:356-383)."""
    rhs = np.ascontiguousarray(rhs, A.dtype)
    x = np.ascontiguousarray(x, A.dtype)
    r = np.zeros(A.rows, A.dtype)
    getattr(lib(), f"oracle_defect_ell_{_sfx(A.dtype)}")(_ptr(r), _ptr(rhs), _ptr(x), *_ell_args(A))
    return r


def dot(x, y):
    x = np.ascontiguousarray(x)
    y = np.ascontiguousarray(y, x.dtype)
    return getattr(lib(), f"oracle_dot_{_sfx(x.dtype)}")(_ptr(x), _ptr(y), _u64(x.size))


def norm2(x, root=False):
    """Norm<vnt_l_two, root, tags::CPU> (la/norm.hh:290-400)."""
    x = np.ascontiguousarray(x)
    v = getattr(lib(), f"oracle_norm2_{_sfx(x.dtype)}")(_ptr(x), _u64(x.size))
    return float(np.sqrt(v)) if root else v


def scaled_sum2(x, y, b):
    """x += b*y (la/scaled_sum.hh:73-88), in place on a copy."""
    x = np.array(x, copy=True)
    y = np.ascontiguousarray(y, x.dtype)
    f = getattr(lib(), f"oracle_scaled_sum2_{_sfx(x.dtype)}")
    f(_ptr(x), _ptr(y), _scalar(x.dtype, b), _u64(x.size))
    return x


def scaled_sum3(x, y, b):
    """r = x + b*y (la/scaled_sum.hh:91-110)."""
    x = np.ascontiguousarray(x)
    y = np.ascontiguousarray(y, x.dtype)
    r = np.empty_like(x)
    getattr(lib(), f"oracle_scaled_sum3_{_sfx(x.dtype)}")(_ptr(r), _ptr(x), _ptr(y), _scalar(x.dtype, b), _u64(x.size))
    return r


def scaled_sum_v(a, b, c):
    """a += b .* c (la/scaled_sum.hh:177-196)."""
    a = np.array(a, copy=True)
    b = np.ascontiguousarray(b, a.dtype)
    c = np.ascontiguousarray(c, a.dtype)
    getattr(lib(), f"oracle_scaled_sum_v_{_sfx(a.dtype)}")(_ptr(a), _ptr(b), _ptr(c), _u64(a.size))
    return a


# ---------------------------------------------------------------------------
# Q1 operator, transfers, prolongation matrix
# ---------------------------------------------------------------------------
def q1_bands(side, dtype=np.float64, bc=0):
    """FillMatrix<POISSON, DIRICHLET_0 (bc=0) | DIRICHLET_NEUMANN (bc=1)>."""
    N = side * side
    bands = np.zeros(9 * N, dtype)
    getattr(lib(), f"oracle_q1_bands_{_sfx(dtype)}")(_u64(N), _ptr(bands), C.c_int(bc))
    return bands.reshape(9, N)


def banded_to_csr(bands):
    """SparseMatrix(const BandedMatrixQx&) (la/sparse_matrix.hh:221-300)."""
    bands = np.ascontiguousarray(bands)
    N = bands.shape[1]
    f = getattr(lib(), f"oracle_banded_q1_to_csr_{_sfx(bands.dtype)}")
    nnz = f(_u64(N), _ptr(bands), None, None, None)
    rp = np.zeros(N + 1, np.uint64)
    col = np.zeros(nnz, np.uint64)
    val = np.zeros(nnz, bands.dtype)
    f(_u64(N), _ptr(bands), _ptr(rp), _ptr(col), _ptr(val))
    return rp, col, val


def banded_q1(bands, x, rhs=None):
    """Product<tags::CPU>::value(BandedMatrixQx<Q1Type>, x) (la/product.hh:
    371-400), or the defect rhs - A x when rhs is given."""
    bands = np.ascontiguousarray(bands)
    n = bands.shape[1]
    x = np.ascontiguousarray(x, bands.dtype)
    y = np.zeros(n, bands.dtype)
    r = None if rhs is None else np.ascontiguousarray(rhs, bands.dtype)
    getattr(lib(), f"oracle_banded_q1_{_sfx(bands.dtype)}")(_u64(n), _ptr(bands), None if r is None else _ptr(r),
                                                             _ptr(x), _ptr(y))
    return y


def q1_coeff_bands(side, kappa, dtype=np.float64):
    """Variable-coefficient Q1 operator as 9 bands (no reference counterpart;
    bilinear element stiffness per cell, Dirichlet rows identity)."""
    kappa = np.ascontiguousarray(kappa, np.float64)
    assert kappa.size == (side - 1) ** 2
    bands = np.zeros(9 * side * side, dtype)
    getattr(lib(), f"oracle_q1_coeff_bands_{_sfx(dtype)}")(_u64(side), _ptr(kappa), _ptr(bands))
    return bands.reshape(9, side * side)


def q1_coeff_matrix(side, kappa, dtype=np.float64, threads=1) -> Ell:
    rp, col, val = banded_to_csr(q1_coeff_bands(side, kappa, dtype))
    n = side * side
    return ell_from_sm(n, n, rp, col, val, threads)


def coarsen_coeff(kappa, nf, averaging=1):
    """Cells of the next coarser grid: mean of 2x2 fine cells (0 arithmetic,
    1 geometric, 2 harmonic)."""
    kappa = np.ascontiguousarray(kappa, np.float64)
    nc = nf // 2
    out = np.zeros(nc * nc)
    lib().oracle_coarsen_coeff(_ptr(kappa), _u64(nf), _ptr(out), _u64(nc), C.c_int(averaging))
    return out


def q1_matrix(side, dtype=np.float64, threads=1, bc=0) -> Ell:
    rp, col, val = banded_to_csr(q1_bands(side, dtype, bc))
    n = side * side
    return ell_from_sm(n, n, rp, col, val, threads)


def restriction(fine, mask=(2,) * 8):
    fine = np.ascontiguousarray(fine)
    nf = fine.size
    side_f = int(round(np.sqrt(nf)))
    side_c = (side_f - 1) // 2 + 1
    coarse = np.zeros(side_c * side_c, fine.dtype)
    m = np.asarray(mask, np.uint64)
    getattr(lib(), f"oracle_restriction_{_sfx(fine.dtype)}")(
        _ptr(coarse), _u64(coarse.size), _ptr(fine), _u64(nf), _ptr(m))
    return coarse


def prolongation(coarse, mask=(2,) * 8):
    coarse = np.ascontiguousarray(coarse)
    nc = coarse.size
    side_c = int(round(np.sqrt(nc)))
    side_f = 2 * (side_c - 1) + 1
    fine = np.zeros(side_f * side_f, coarse.dtype)
    m = np.asarray(mask, np.uint64)
    getattr(lib(), f"oracle_prolongation_{_sfx(coarse.dtype)}")(
        _ptr(fine), _u64(fine.size), _ptr(coarse), _u64(nc), _ptr(m))
    return fine


def apply_dirichlet(grid, mask=(2,) * 8):
    g = np.array(grid, copy=True)
    m = np.asarray(mask, np.uint64)
    getattr(lib(), f"oracle_apply_dirichlet_{_sfx(g.dtype)}")(_ptr(g), _u64(g.size), _ptr(m))
    return g


def prolmat_csr(fine_side):
    """ProlongationMatrix (math/prolongation_matrix.hh:28-100), natural order."""
    f = lib().oracle_prolmat_csr
    nnz = f(_u64(fine_side), None, None, None)
    rp = np.zeros(fine_side * fine_side + 1, np.uint64)
    col = np.zeros(nnz, np.uint64)
    val = np.zeros(nnz, np.float64)
    f(_u64(fine_side), _ptr(rp), _ptr(col), _ptr(val))
    return rp, col, val


def prolmat(fine_side, dtype=np.float64, threads=1) -> Ell:
    rp, col, val = prolmat_csr(fine_side)
    side_c = (fine_side + 1) // 2
    return ell_from_sm(fine_side * fine_side, side_c * side_c, rp, col, val.astype(dtype), threads)


def transpose(A: Ell, threads=None) -> Ell:
    """Transposition<Tag> (math/transposition.hh:31-50) via SparseMatrix."""
    rp, col, val = A.to_csr()
    rows = A.cols
    entries = [[] for _ in range(rows)]
    for r in range(A.rows):
        for k in range(int(rp[r]), int(rp[r + 1])):
            entries[int(col[k])].append((r, val[k]))
    trp = [0]
    tcol, tval = [], []
    for r in range(rows):
        for c, v in sorted(entries[r]):
            tcol.append(c)
            tval.append(v)
        trp.append(len(tcol))
    return ell_from_sm(rows, A.rows, np.asarray(trp, np.uint64), np.asarray(tcol, np.uint64),
                       np.asarray(tval, A.dtype), A.threads if threads is None else threads)


# ---------------------------------------------------------------------------
# Solvers
# ---------------------------------------------------------------------------
def cg_solver(A: Ell, b, x0=None, max_iters=10000, eps=1e-8, history=False):
    """CGSolver<tags::CPU, NONE> (math/cg.hh:63-133)."""
    dt = A.dtype
    b = np.ascontiguousarray(b, dt)
    x = np.zeros(A.rows, dt) if x0 is None else np.array(x0, dt, copy=True)
    used = _u64()
    hist = np.zeros(max_iters, dt) if history else None
    getattr(lib(), f"oracle_cg_solver_{_sfx(dt)}")(
        *_ell_args(A), _ptr(b), _ptr(x), _u64(max_iters), _scalar(dt, eps), C.byref(used),
        _ptr(hist) if history else None)
    if history:
        return x, used.value, hist[:used.value]
    return x, used.value


def pcg_jacobi(A: Ell, b, dinv, x0=None, max_iters=10000, eps=1e-8, history=False):
    """ConjugateGradients<tags::CPU, JAC>::value(A, b, x, dd_inverted,
    max_iters, used_iters, eps_relative) on ELL
    (math/conjugate_gradients.hh:1110-1154).  used_iters is 0 when the
    reference leaves it unset (max_iters = 0)."""
    dt = A.dtype
    b = np.ascontiguousarray(b, dt)
    dinv = np.ascontiguousarray(dinv, dt)
    x = np.zeros(A.rows, dt) if x0 is None else np.array(x0, dt, copy=True)
    used = _u64(0)
    hist = np.zeros(max(max_iters, 1), dt) if history else None
    getattr(lib(), f"oracle_pcg_jacobi_{_sfx(dt)}")(
        *_ell_args(A), _ptr(b), _ptr(x), _ptr(dinv), _u64(max_iters), _scalar(dt, eps), C.byref(used),
        _ptr(hist) if history else None)
    if history:
        return x, used.value, hist[:used.value]
    return x, used.value


def csr_arrays(A: Ell):
    """Row pointers, columns and values of A's stored non-zeros, rows in
    ascending column order: the entries of Ell.to_csr
    (la/sparse_matrix.hh:157-185), vectorised for large matrices.  Slot
    s = r*T + t + k*stride belongs to row (s mod stride) / T and counts while
    k < Arl[r]; explicit zeros are dropped as SparseMatrix drops them."""
    T, S = A.threads, A.stride
    nch = len(A.Ax) // S if S else 0
    lanes = (np.arange(A.rows, dtype=np.int64)[:, None] * T + np.arange(T, dtype=np.int64)[None, :]).reshape(-1)
    rows = np.repeat(np.arange(A.rows, dtype=np.int64), T)
    arl = A.Arl.astype(np.int64)[rows]
    k = np.arange(nch, dtype=np.int64)[:, None]
    slots = (lanes[None, :] + k * S)[k < arl[None, :]]
    rr = np.broadcast_to(rows[None, :], (nch, rows.size))[k < arl[None, :]]
    vals = A.Ax[slots]
    keep = vals != 0
    slots, rr, vals = slots[keep], rr[keep], vals[keep]
    cols = A.Aj[slots].astype(np.int64)
    order = np.lexsort((cols, rr))
    rp = np.zeros(A.rows + 1, np.int64)
    np.add.at(rp, rr + 1, 1)
    return np.cumsum(rp), cols[order], vals[order]


def direct_solve(A: Ell, b, tol=1e-14, max_refine=4):
    """FP64 reference solution of A x = b (no reference counterpart): a
    sparse LU factorisation (SuperLU 5 through scipy.sparse.linalg.splu, COLAMD
    column ordering), then iterative refinement x += LU^-1 (b - A x) with the
    residual formed by the oracle's Defect<tags::CPU> (math/defect.hh:449-504)
    until ||b - A x|| <= tol ||b|| or max_refine steps.  Returns (x, relative
    residual, refinement steps).  Pinned by closed forms (b = A x_true gives
    back x_true), by agreement with CGSolver run to a tight tolerance and by
    its own residual (tests/test_oracle.py)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    A64 = A if A.dtype == np.float64 else A.astype(np.float64)
    b = np.ascontiguousarray(b, np.float64)
    rp, cols, vals = csr_arrays(A64)
    M = sp.csc_matrix(sp.csr_matrix((vals, cols, rp), shape=(A64.rows, A64.cols)))
    lu = spl.splu(M, permc_spec="COLAMD")
    x = lu.solve(b)
    nb = float(np.linalg.norm(b)) or 1.0
    rel, steps = float(np.linalg.norm(defect(A64, b, x))) / nb, 0
    while rel > tol and steps < max_refine:
        x = x + lu.solve(defect(A64, b, x))
        rel = float(np.linalg.norm(defect(A64, b, x))) / nb
        steps += 1
    return x, rel, steps


def cg_konv(A: Ell, b, konv=None):
    """ConjugateGradients<tags::CPU, NONE>::value(A, b, x, konv_rad)."""
    dt = A.dtype
    b = np.ascontiguousarray(b, dt)
    x = np.zeros(A.rows, dt)
    if konv is None:
        konv = float(np.finfo(dt).eps)
    it = getattr(lib(), f"oracle_cg_konv_{_sfx(dt)}")(*_ell_args(A), _ptr(b), _ptr(x), C.c_double(konv))
    return x, it


class OracleEll(C.Structure):
    _fields_ = [("rows", _u64), ("cols", _u64), ("stride", _u64), ("ncpr", _u64), ("threads", _u64),
                ("Aj", _p), ("Ax", _p), ("Arl", _p)]


class OracleMGLevel(C.Structure):
    _fields_ = [("A", OracleEll), ("dinv", _p), ("prol", OracleEll), ("res", OracleEll)]


class OracleMGConfig(C.Structure):
    _fields_ = [("min_level", _u64), ("max_level", _u64), ("n_max_iter", _u64),
                ("n_pre_smooth", _u64), ("n_post_smooth", _u64),
                ("initial_zero", C.c_int), ("convergence_check", C.c_int),
                ("is_smoother", C.c_int), ("prolmat", C.c_int),
                ("tolerance", C.c_double), ("adapt_correction_factor", C.c_double),
                ("mask", _u64 * 8), ("zero_inner_start", C.c_int), ("max_outer", _u64)]


@dataclass
class MGHierarchy:
    """Per-level data of MGInfo (math/multigrid.hh:66-149); index = level."""
    A: dict = field(default_factory=dict)
    dinv: dict = field(default_factory=dict)
    prol: dict = field(default_factory=dict)
    res: dict = field(default_factory=dict)

    def levels_struct(self, max_level):
        arr = (OracleMGLevel * (max_level + 1))()
        for l in range(max_level + 1):
            if l in self.A:
                arr[l].A = self.A[l].c_struct()
                arr[l].dinv = _ptr(self.dinv[l])
            if l in self.prol:
                arr[l].prol = self.prol[l].c_struct()
                arr[l].res = self.res[l].c_struct()
        return arr


def q1_coeff_hierarchy(max_level, kappa, averaging=1, omega=0.8, dtype=np.float32, min_level=1):
    """Rediscretised variable-coefficient hierarchy: level l has (2^l+1)^2
    nodes and the coefficient coarsened level by level; dinv = omega/diag."""
    H = MGHierarchy()
    k = np.ascontiguousarray(kappa, np.float64)
    for l in range(max_level, min_level - 1, -1):
        side = 2 ** l + 1
        bands = q1_coeff_bands(side, k, dtype)
        rp, col, val = banded_to_csr(bands)
        n = side * side
        H.A[l] = ell_from_sm(n, n, rp, col, val)
        d = bands[4].astype(dtype)
        H.dinv[l] = ((np.ones(n, dtype) / d).astype(dtype) * dtype(omega)).astype(dtype)
        if l > min_level:
            k = coarsen_coeff(k, 2 ** l, averaging)
    return H


def q1_hierarchy(max_level, min_level=1, dtype=np.float32, omega=0.7, use_prolmat=False, threads=1):
    """Geometric Q1 hierarchy as in math/poisson_mg_*_ell_TEST.cc: level l has
    (2^l+1)^2 nodes; diag_inverted = omega / diag(A)."""
    H = MGHierarchy()
    for l in range(min_level, max_level + 1):
        side = 2 ** l + 1
        bands = q1_bands(side, dtype)
        rp, col, val = banded_to_csr(bands)
        n = side * side
        H.A[l] = ell_from_sm(n, n, rp, col, val, threads)
        d = bands[4].astype(dtype)
        inv = (np.ones(n, dtype) / d).astype(dtype)
        H.dinv[l] = (inv * dtype(omega)).astype(dtype)
        if use_prolmat and l > min_level:
            P = prolmat(side, dtype, threads)
            H.prol[l] = P
            H.res[l] = transpose(P)
    return H


def mg_config(min_level, max_level, n_max_iter=4, n_pre=2, n_post=2, tolerance=1e-8,
              initial_zero=False, convergence_check=False, adapt=1.0, prolmat=False,
              mask=(2,) * 8, zero_inner_start=False, max_outer=0):
    c = OracleMGConfig()
    c.min_level, c.max_level, c.n_max_iter = min_level, max_level, n_max_iter
    c.n_pre_smooth, c.n_post_smooth = n_pre, n_post
    c.initial_zero, c.convergence_check, c.is_smoother = int(initial_zero), int(convergence_check), 0
    c.prolmat = int(prolmat)
    c.tolerance, c.adapt_correction_factor = tolerance, adapt
    for i in range(8):
        c.mask[i] = mask[i]
    c.zero_inner_start = int(zero_inner_start)
    c.max_outer = max_outer
    return c


def mg_fixed(H: MGHierarchy, cfg, b, x0=None):
    """Multigrid<tags::CPU, tags::CPU, NONE|PROLMAT, JAC, V, FIXED>::value."""
    lv = H.levels_struct(cfg.max_level)
    dt = H.A[cfg.max_level].dtype
    b = np.ascontiguousarray(b, dt)
    x = np.zeros(b.size, dt) if x0 is None else np.array(x0, dt, copy=True)
    fd = C.c_double()
    it = getattr(lib(), f"oracle_mg_fixed_{_sfx(dt)}")(lv, C.byref(cfg), _ptr(b), _ptr(x), C.byref(fd))
    return x, it, fd.value


def mg_mixed(H: MGHierarchy, cfg, A64: Ell, b, x0=None):
    """Multigrid<..., JAC, V, MIXED>::value (math/multigrid.hh:2395-2536).
    Returns (x, outer_iterations, inner_cycles, relative_defect_history)."""
    lv = H.levels_struct(cfg.max_level)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(b.size) if x0 is None else np.array(x0, np.float64, copy=True)
    inner = _u64()
    cap = cfg.max_outer or 1000
    hist = np.zeros(cap)
    a = A64.c_struct()
    outer = lib().oracle_mg_mixed(lv, C.byref(cfg), C.byref(a), _ptr(b), _ptr(x), C.byref(inner), _ptr(hist))
    return x, outer, inner.value, hist[:min(outer, cap)]


def mg_oplist(H: MGHierarchy, min_level, max_level, b, x0, cycle="W", max_iters=100,
              max_iters_coarse=100, n_pre=4, n_post=4, eps=1e-8):
    """MGSolver over MGCycleCreation<cycle, CGSolver, RISmoother, PROLMAT>."""
    lv = H.levels_struct(max_level)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, np.float64, copy=True)
    norms = np.zeros(2)
    used = lib().oracle_mg_oplist_d(lv, _u64(min_level), _u64(max_level),
                                    C.c_int({"V": 0, "W": 1, "F": 2}[cycle]), _u64(max_iters),
                                    _u64(max_iters_coarse), _u64(n_pre), _u64(n_post),
                                    C.c_double(eps), _ptr(b), _ptr(x), _ptr(norms))
    return x, used, norms


# ---------------------------------------------------------------------------
# File formats of the reference (numpy restatements)
# ---------------------------------------------------------------------------
def read_ell(path, dtype=np.float64) -> Ell:
    """MatrixIO<io_formats::ELL>::read_matrix (math/matrix_io.hh:466-584):
    5 uint64 header words (size, rows, columns, stride, num_cols_per_row),
    uint64 Aj[size], float64 Ax[size]; threads = 1; Arl from Ax."""
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:40].view(np.uint64)
    size, rows, cols, stride, ncpr = (int(v) for v in hdr)
    Aj = raw[40:40 + 8 * size].view(np.uint64).copy()
    Ax = raw[40 + 8 * size:40 + 16 * size].view(np.float64).astype(dtype)
    return ell_from_arrays(rows, cols, stride, ncpr, Aj, Ax, 1)


def write_ell(path, A: Ell):
    """MatrixIO<io_formats::ELL>::write_matrix (math/matrix_io.hh:436-464)."""
    if A.threads != 1:
        raise ValueError("Only Matrices with 1 threads data layout are supported for export")
    hdr = np.array([A.Aj.size, A.rows, A.cols, A.stride, A.ncpr], np.uint64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(A.Aj.astype(np.uint64).tobytes())
        f.write(A.Ax.astype(np.float64).tobytes())


def read_exp(path, dtype=np.float64):
    """VectorIO<io_formats::EXP>::read_vector (math/vector_io.hh:52-115):
    lines containing '#' are skipped, one value per line."""
    vals = []
    with open(path) as f:
        for line in f:
            if "#" in line:
                continue
            s = line.strip()
            if not s:
                continue
            vals.append(float(s))
    return np.asarray(vals, dtype)


def read_dv(path, dtype=np.float64):
    """VectorIO<io_formats::DV>::read_vector (math/vector_io.hh:226-246)."""
    raw = np.fromfile(path, dtype=np.uint8)
    n = int(raw[:8].view(np.uint64)[0])
    return raw[8:8 + 8 * n].view(np.float64).astype(dtype)


def read_mtx(path):
    """MatrixMarket coordinate file -> (rows, cols, row-sorted COO) with
    duplicate-free entries, as the reference's MTX reader fills a
    SparseMatrix (clients/io_tools/mtxtoell.cc:46-50)."""
    with open(path) as f:
        lines = [l for l in f if not l.startswith("%")]
    rows, cols, nnz = (int(v) for v in lines[0].split()[:3])
    ent = {}
    for l in lines[1:1 + nnz]:
        p = l.split()
        if len(p) < 3:
            continue
        ent[(int(p[0]) - 1, int(p[1]) - 1)] = float(p[2])
    keys = sorted(ent)
    rp = np.zeros(rows + 1, np.uint64)
    col = np.array([k[1] for k in keys], np.uint64)
    val = np.array([ent[k] for k in keys], np.float64)
    for (r, _c) in keys:
        rp[r + 1] += 1
    rp = np.cumsum(rp).astype(np.uint64)
    return rows, cols, rp, col, val
