"""honei_amd -- MI355X (gfx950) backend for HONEI's ELL sparse-solve path.

The product is the shared library ``honei_amd/lib/libhonei_hip.so`` (HIP
kernels + C++ host code) behind the C ABI in ``include/honei_hip.h``.  This
module is a thin ctypes binding used by the tests and the benchmark; it
raises immediately when the library is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libhonei_hip.so")

F32, F64, BF16 = 0, 1, 2
SUCCESS = 0
ERRORS = {1: "INVALID_ARGUMENT", 2: "SIZE_MISMATCH", 3: "OUT_OF_MEMORY", 4: "RUNTIME",
          5: "NOT_CONVERGED", 6: "BUDGET", 7: "NO_DEVICE"}

_sz = C.c_size_t
_p = C.c_void_p
_i = C.c_int
_d = C.c_double
_u64 = C.c_uint64

# name -> argtypes (restype is int unless listed in _RESTYPE)
_SIGS = {
    "honei_hip_device_count": [C.POINTER(_i)],
    "honei_hip_init": [_i],
    "honei_hip_shutdown": [],
    "honei_hip_malloc": [C.POINTER(_p), _sz],
    "honei_hip_host_alloc": [C.POINTER(_p), _sz],
    "honei_hip_host_free": [_p],
    "honei_hip_free": [_p],
    "honei_hip_upload": [_p, _p, _sz, _p],
    "honei_hip_download": [_p, _p, _sz, _p],
    "honei_hip_copy": [_p, _p, _sz, _p],
    "honei_hip_fill_zero": [_p, _sz, _p],
    "honei_hip_fill": [_i, _p, _d, _sz, _p],
    "honei_hip_synchronize": [_p],
    "honei_hip_stream_create": [C.POINTER(_p)],
    "honei_hip_stream_destroy": [_p],
    "honei_hip_event_create": [C.POINTER(_p)],
    "honei_hip_event_destroy": [_p],
    "honei_hip_event_record": [_p, _p],
    "honei_hip_event_elapsed_ms": [C.POINTER(C.c_float), _p, _p],
    "honei_hip_convert_double_float": [_p, _p, _sz, _p],
    "honei_hip_convert_float_double": [_p, _p, _sz, _p],
    "honei_hip_upload_indices": [_p, _p, _sz, _u64],
    "honei_hip_upload_values_bf16": [_p, _p, _sz],
    "honei_hip_product_smell_dv_float": [_p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_product_smell_dv_double": [_p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_product_smell_dv_bf16": [_p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_defect_smell_dv_float": [_p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_defect_smell_dv_double": [_p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_product_smell_dv_float_idx64": [_p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_product_smell_dv_double_idx64": [_p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_defect_smell_dv_float_idx64": [_p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_defect_smell_dv_double_idx64": [_p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_narrow_indices": [_p, _p, _sz, _u64, _p],
    "honei_hip_stream_probe": [_i, _p, _p, _sz, _sz, _p],
    "honei_hip_sell_create": [C.POINTER(_p), _i, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _p],
    "honei_hip_sell_info": [_p, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz)],
    "honei_hip_product_sell": [_p, _p, _p, _p],
    "honei_hip_sell_destroy": [_p],
    "honei_hip_sell_stage_x": [_p, _i],
    "honei_hip_defect_norm_smell_dv": [_i, _p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p],
    "honei_hip_jacobi_smell_dv": [_i, _p, _p, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p],
    "honei_hip_dot_product_two_float": [C.POINTER(C.c_float), _p, _p, _sz, _p],
    "honei_hip_dot_product_two_double": [C.POINTER(_d), _p, _p, _sz, _p],
    "honei_hip_norm_l2_one_float": [C.POINTER(C.c_float), _p, _sz, _p],
    "honei_hip_norm_l2_one_double": [C.POINTER(_d), _p, _sz, _p],
    "honei_hip_dot_product_dev": [_i, _p, _p, _p, _sz, _p],
    "honei_hip_norm_l2_dev": [_i, _p, _p, _sz, _p],
    "honei_hip_axpy_norm_dev": [_i, _p, _p, _d, _p, _sz, _p],
    "honei_hip_scaled_sum_two": [_i, _p, _p, _d, _sz, _p],
    "honei_hip_scaled_sum_three_s": [_i, _p, _p, _p, _d, _sz, _p],
    "honei_hip_scaled_sum_three": [_i, _p, _p, _p, _sz, _p],
    "honei_hip_sum_two": [_i, _p, _p, _sz, _p],
    "honei_hip_difference_two": [_i, _p, _p, _sz, _p],
    "honei_hip_difference_three": [_i, _p, _p, _p, _sz, _p],
    "honei_hip_scale_one": [_i, _p, _d, _sz, _p],
    "honei_hip_element_product_three": [_i, _p, _p, _p, _sz, _p],
    "honei_hip_element_inverse_one": [_i, _p, _sz, _p],
    "honei_hip_restriction": [_i, _p, _sz, _p, _sz, _p, _p],
    "honei_hip_prolongation": [_i, _p, _sz, _p, _sz, _p, _p],
    "honei_hip_q1_assemble": [_i, _sz, _p, _p, _p, _sz, _p, _p, _d, _sz, _p],
    "honei_hip_random_ell": [_i, _sz, _sz, _sz, _u64, _p, _p, _p, _sz, _sz, _p],
    "honei_hip_cg_solve": [_i, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p, _sz, _d, C.POINTER(_sz),
                           C.POINTER(_d), _p],
    "honei_hip_cg_solve_noalloc": [_i, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p, _sz, _d, C.POINTER(_sz),
                                   C.POINTER(_d), _p],
    "honei_hip_cg_konv": [_i, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p, _d, _sz, C.POINTER(_sz), _p],
    "honei_hip_pcg_jacobi": [_i, _p, _p, _p, _sz, _sz, _sz, _sz, _p, _p, _p, _sz, _d, C.POINTER(_sz),
                             C.POINTER(_d), _p],
    "honei_hip_mg_create": [C.POINTER(_p), _i, _sz, _sz],
    "honei_hip_mg_destroy": [_p],
    "honei_hip_mg_configure": [_p, _p],
    "honei_hip_mg_set_level": [_p, _sz, _sz, _sz, _sz, _sz, _p, _p, _p, _p],
    "honei_hip_mg_set_transfer": [_p, _sz, _i, _sz, _sz, _sz, _sz, _sz, _p, _p, _p],
    "honei_hip_mg_set_outer": [_p, _sz, _sz, _sz, _sz, _p, _p, _p],
    "honei_hip_mg_build_q1": [_p, _d, _p],
    "honei_hip_mg_build_q1_coeff": [_p, _d, _p, _i, _p],
    "honei_hip_mg_solve_fixed": [_p, _p, _p, C.POINTER(_sz), C.POINTER(_d), _p],
    "honei_hip_mg_solve_mixed": [_p, _p, _p, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_d), _p],
    "honei_hip_mg_solve_pcg": [_p, _p, _p, C.POINTER(_sz), C.POINTER(_d), _p],
    "honei_hip_product_bmdv_q1_float": [_p] * 11 + [_sz, _sz, _sz, _p],
    "honei_hip_product_bmdv_q1_double": [_p] * 11 + [_sz, _sz, _sz, _p],
    "honei_hip_defect_q1_float": [_p] * 12 + [_sz, _sz, _sz, _p],
    "honei_hip_defect_q1_double": [_p] * 12 + [_sz, _sz, _sz, _p],
    "honei_hip_fill_uniform": [_p, _sz, _d, _d, _u64, _p],
    "honei_hip_fill_lognormal": [_p, _sz, _d, _u64, _p],
    "honei_hip_batch_config_default": [_p],
    "honei_hip_batch_create": [C.POINTER(_p), _p, _p, C.c_int32, _p, _u64],
    "honei_hip_batch_destroy": [_p],
    "honei_hip_batch_info": [_p, C.POINTER(C.c_int64), C.POINTER(_u64), C.POINTER(C.c_int64)],
    "honei_hip_batch_slot": [_p, C.c_int32, C.c_int64, C.POINTER(C.c_int64)],
    "honei_hip_batch_pair_timing": [_p, C.c_int32, C.POINTER(_d), C.POINTER(_d)],
    "honei_hip_batch_time_level": [_p, C.c_int32, C.c_int32, C.POINTER(_d)],
    "honei_hip_batch_load": [_p, C.c_int64, _p, _p],
    "honei_hip_batch_load_many": [_p, C.c_int64, _p, _p, _p],
    "honei_hip_batch_load_seeded": [_p, C.c_int64, _p, _p, _p, _d],
    "honei_hip_batch_activate": [_p, C.c_int32],
    "honei_hip_batch_step": [_p],
    "honei_hip_batch_poll": [_p, _p],
    "honei_hip_batch_result": [_p, C.c_int64, _p],
    "honei_hip_batch_run": [_p, C.POINTER(C.c_int64)],
    "honei_hip_batch_shape": [_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                              C.POINTER(_d), C.POINTER(C.c_int64)],
    "honei_hip_serve_run": [_p, _p, C.c_int64, _p, _p],
    "honei_hip_ell_block_ranges": [_p, _p, _sz, _sz, _sz, _sz, _sz, _p, C.POINTER(C.c_int64), _p],
    "honei_hip_product_smell_dv_staged": [_i, _p, _p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz, _sz, _p, _sz, C.c_int64,
                                          _p],
}
_RESTYPE = {"honei_hip_last_error": C.c_char_p, "honei_hip_version": C.c_char_p}

_lib = None


class HipError(RuntimeError):
    def __init__(self, code, fn, msg):
        super().__init__(f"{fn} failed: {ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Load the backend library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP backend library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _i
        for name, rt in _RESTYPE.items():
            f = getattr(L, name)
            f.argtypes = []
            f.restype = rt
        _lib = L
    return _lib


def declared_symbols():
    return list(_SIGS) + list(_RESTYPE)


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != SUCCESS:
        raise HipError(rc, name, lib().honei_hip_last_error().decode())
    return rc


def last_error() -> str:
    return lib().honei_hip_last_error().decode()


def init(device: int = 0):
    call("honei_hip_init", device)


def device_count() -> int:
    n = _i(0)
    rc = lib().honei_hip_device_count(C.byref(n))
    return n.value if rc == SUCCESS else 0


def synchronize(stream=None):
    call("honei_hip_synchronize", stream)


class Event:
    def __init__(self):
        p = _p()
        call("honei_hip_event_create", C.byref(p))
        self.ptr = p.value

    def record(self, stream=None):
        call("honei_hip_event_record", self.ptr, stream)

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        call("honei_hip_event_elapsed_ms", C.byref(ms), self.ptr, stop.ptr)
        return ms.value

    def __del__(self):
        try:
            lib().honei_hip_event_destroy(self.ptr)
        except Exception:
            pass


class DeviceArray:
    """Owning device buffer with a numpy dtype (test and bench plumbing)."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = _p()
        call("honei_hip_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, a.dtype)
        if a.size:
            call("honei_hip_upload", d.ptr, a.ctypes.data_as(_p), d.nbytes, None)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        call("honei_hip_upload", self.ptr, a.ctypes.data_as(_p), self.nbytes, None)

    def get(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.n:
            call("honei_hip_download", out.ctypes.data_as(_p), self.ptr, self.nbytes, None)
        return out

    def zero(self, stream=None):
        call("honei_hip_fill_zero", self.ptr, self.nbytes, stream)

    def free(self):
        if self.ptr:
            lib().honei_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceEll:
    """ELL matrix resident in HBM: int32 Aj, uint32 Arl, values f32/f64/bf16.

    Built from the reference's host layout (uint64 Aj/Arl, same stride and
    slot order); indices are narrowed and range-checked on upload."""

    def __init__(self, rows, cols, stride, ncpr, threads, Aj, Ax, Arl, vtype=None):
        self.rows, self.cols, self.stride, self.ncpr, self.threads = (int(rows), int(cols), int(stride),
                                                                      int(ncpr), int(threads))
        Ax = np.ascontiguousarray(Ax)
        if vtype is None:
            vtype = F64 if Ax.dtype == np.float64 else F32
        self.vtype = vtype
        n = Ax.size
        self.Aj = DeviceArray(n, np.int32)
        call("honei_hip_upload_indices", self.Aj.ptr, np.ascontiguousarray(Aj, np.uint64).ctypes.data_as(_p),
             n, max(self.cols, 1))
        self.Arl = DeviceArray(self.rows, np.uint32)
        call("honei_hip_upload_indices", self.Arl.ptr, np.ascontiguousarray(Arl, np.uint64).ctypes.data_as(_p),
             self.rows, self.ncpr + 1)
        if vtype == BF16:
            self.Ax = DeviceArray(n, np.uint16)
            call("honei_hip_upload_values_bf16", self.Ax.ptr,
                 np.ascontiguousarray(Ax, np.float32).ctypes.data_as(_p), n)
        else:
            self.Ax = DeviceArray.from_host(Ax, np.float64 if vtype == F64 else np.float32)

    @classmethod
    def from_oracle(cls, A, vtype=None):
        return cls(A.rows, A.cols, A.stride, A.ncpr, A.threads, A.Aj, A.Ax, A.Arl, vtype)

    @property
    def vec_dtype(self):
        return np.float64 if self.vtype == F64 else np.float32

    def product(self, y: DeviceArray, x: DeviceArray, row_start=0, row_end=None, blocksize=0, stream=None):
        fn = {F32: "honei_hip_product_smell_dv_float", F64: "honei_hip_product_smell_dv_double",
              BF16: "honei_hip_product_smell_dv_bf16"}[self.vtype]
        call(fn, x.ptr, y.ptr, self.Aj.ptr, self.Ax.ptr, self.Arl.ptr, row_start,
             self.rows if row_end is None else row_end, self.ncpr, self.stride, blocksize, self.threads, stream)

    def block_ranges(self, rows_per_block, stream=None):
        """(device int32 [nblocks][2] column windows, widest window)."""
        nb = (self.rows + rows_per_block - 1) // rows_per_block
        ranges = DeviceArray(max(2 * nb, 2), np.int32)
        w = C.c_int64()
        call("honei_hip_ell_block_ranges", self.Aj.ptr, self.Arl.ptr, self.rows, self.ncpr, self.stride,
             self.threads, rows_per_block, ranges.ptr, C.byref(w), stream)
        return ranges, w.value

    def product_staged(self, y: DeviceArray, x: DeviceArray, ranges: DeviceArray, window, blocksize=256, chunks=1,
                       stream=None):
        """y = A x with each work-group's x window in LDS (ranges made with
        rows_per_block = chunks * blocksize / threads)."""
        call("honei_hip_product_smell_dv_staged", self.vtype, x.ptr, y.ptr, self.Aj.ptr, self.Ax.ptr, self.Arl.ptr,
             self.rows, self.ncpr, self.stride, self.threads, blocksize, chunks, ranges.ptr, ranges.n // 2,
             int(window), stream)

    def defect(self, r: DeviceArray, rhs: DeviceArray, x: DeviceArray, stream=None):
        fn = {F32: "honei_hip_defect_smell_dv_float", F64: "honei_hip_defect_smell_dv_double"}[self.vtype]
        call(fn, rhs.ptr, r.ptr, self.Aj.ptr, self.Ax.ptr, self.Arl.ptr, x.ptr, 0, self.rows, self.ncpr,
             self.stride, 0, self.threads, stream)

    def defect_norm(self, r, rhs, x, norm_dev: DeviceArray, stream=None):
        call("honei_hip_defect_norm_smell_dv", self.vtype, rhs.ptr, r.ptr, self.Aj.ptr, self.Ax.ptr, self.Arl.ptr,
             x.ptr, self.rows, self.ncpr, self.stride, self.threads, norm_dev.ptr, stream)

    def jacobi(self, x_new, x, rhs, dinv, acc=None, stream=None):
        call("honei_hip_jacobi_smell_dv", self.vtype, x_new.ptr, None if x is None else x.ptr, rhs.ptr, dinv.ptr,
             self.Aj.ptr, self.Ax.ptr, self.Arl.ptr, self.rows, self.ncpr, self.stride, self.threads,
             None if acc is None else acc.ptr, stream)


class Sell:
    """Sliced copy of a threads = 1 DeviceEll (honei_hip_sell_*): rows sorted
    by length within windows of 2048, slices of 64 padded to their longest
    row; products bit-identical to the ELL product."""

    def __init__(self, D: "DeviceEll", stream=None):
        p = _p()
        call("honei_hip_sell_create", C.byref(p), D.vtype, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, D.rows, D.cols, D.ncpr,
             D.stride, D.threads, stream)
        self.h = p.value
        self.rows = D.rows
        sl, pad, nb = _sz(), _sz(), _sz()
        call("honei_hip_sell_info", self.h, C.byref(sl), C.byref(pad), C.byref(nb))
        self.slices, self.padded_slots, self.device_bytes = sl.value, pad.value, nb.value

    def product(self, y: "DeviceArray", x: "DeviceArray", stream=None):
        call("honei_hip_product_sell", self.h, x.ptr, y.ptr, stream)

    def stage_x(self, on: bool):
        """Gather x from an LDS copy of each window's column range (default)
        or from global memory; the same bits either way."""
        call("honei_hip_sell_stage_x", self.h, 1 if on else 0)

    def close(self):
        if getattr(self, "h", None):
            lib().honei_hip_sell_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MGConfig(C.Structure):
    """honei_hip_mg_config (include/honei_hip.h)."""
    _fields_ = [("n_max_iter", _sz), ("n_pre_smooth", _sz), ("n_post_smooth", _sz),
                ("initial_zero", _i), ("convergence_check", _i), ("prolmat", _i), ("zero_inner_start", _i),
                ("tolerance", _d), ("adapt_correction_factor", _d), ("mask", _u64 * 8), ("max_outer", _sz),
                ("post_smooth_uses_pre", _i)]


def mg_config(n_max_iter=4, n_pre=2, n_post=2, tolerance=1e-8, initial_zero=False, adapt=1.0,
              prolmat=False, zero_inner_start=False, mask=(2,) * 8, max_outer=0, post_smooth_uses_pre=True):
    c = MGConfig()
    c.n_max_iter, c.n_pre_smooth, c.n_post_smooth = n_max_iter, n_pre, n_post
    c.initial_zero, c.convergence_check, c.prolmat = int(initial_zero), 0, int(prolmat)
    c.zero_inner_start = int(zero_inner_start)
    c.tolerance, c.adapt_correction_factor = tolerance, adapt
    for i in range(8):
        c.mask[i] = mask[i]
    c.max_outer = max_outer
    c.post_smooth_uses_pre = int(post_smooth_uses_pre)
    return c


class Multigrid:
    """Owning handle to a device-resident multigrid hierarchy."""

    def __init__(self, inner_vtype, min_level, max_level):
        p = _p()
        call("honei_hip_mg_create", C.byref(p), inner_vtype, min_level, max_level)
        self.h = p.value
        self.inner = inner_vtype
        self.min_level, self.max_level = min_level, max_level

    def configure(self, cfg: MGConfig):
        call("honei_hip_mg_configure", self.h, C.byref(cfg))

    def build_q1(self, omega=0.7, stream=None):
        call("honei_hip_mg_build_q1", self.h, omega, stream)

    def build_q1_coeff(self, coeff: "DeviceArray", omega=0.7, averaging=0, stream=None):
        call("honei_hip_mg_build_q1_coeff", self.h, omega, coeff.ptr, averaging, stream)

    def set_level(self, level, A, dinv):
        Ax = np.ascontiguousarray(A.Ax, np.float64 if self.inner == F64 else np.float32)
        dv = np.ascontiguousarray(dinv, np.float64 if self.inner == F64 else np.float32)
        call("honei_hip_mg_set_level", self.h, level, A.rows, A.stride, A.ncpr, A.threads,
             np.ascontiguousarray(A.Aj, np.uint64).ctypes.data_as(_p), Ax.ctypes.data_as(_p),
             np.ascontiguousarray(A.Arl, np.uint64).ctypes.data_as(_p), dv.ctypes.data_as(_p))

    def set_transfer(self, level, kind, A):
        Ax = np.ascontiguousarray(A.Ax, np.float64 if self.inner == F64 else np.float32)
        call("honei_hip_mg_set_transfer", self.h, level, kind, A.rows, A.cols, A.stride, A.ncpr, A.threads,
             np.ascontiguousarray(A.Aj, np.uint64).ctypes.data_as(_p), Ax.ctypes.data_as(_p),
             np.ascontiguousarray(A.Arl, np.uint64).ctypes.data_as(_p))

    def set_outer(self, A):
        Ax = np.ascontiguousarray(A.Ax, np.float64)
        call("honei_hip_mg_set_outer", self.h, A.rows, A.stride, A.ncpr, A.threads,
             np.ascontiguousarray(A.Aj, np.uint64).ctypes.data_as(_p), Ax.ctypes.data_as(_p),
             np.ascontiguousarray(A.Arl, np.uint64).ctypes.data_as(_p))

    def solve_fixed(self, b: DeviceArray, x: DeviceArray, stream=None):
        it, fd = _sz(), _d()
        call("honei_hip_mg_solve_fixed", self.h, b.ptr, x.ptr, C.byref(it), C.byref(fd), stream)
        return it.value, fd.value

    def solve_pcg(self, b: DeviceArray, x: DeviceArray, stream=None, check=True):
        """FP64 CG preconditioned by one V-cycle per iteration; (iterations, rel. defect)."""
        i, rd = _sz(), _d()
        rc = lib().honei_hip_mg_solve_pcg(self.h, b.ptr, x.ptr, C.byref(i), C.byref(rd), stream)
        if rc != SUCCESS and (check or rc != 5):
            raise HipError(rc, "honei_hip_mg_solve_pcg", last_error())
        return i.value, rd.value

    def solve_mixed(self, b: DeviceArray, x: DeviceArray, stream=None, check=True):
        o, i, rd = _sz(), _sz(), _d()
        rc = lib().honei_hip_mg_solve_mixed(self.h, b.ptr, x.ptr, C.byref(o), C.byref(i), C.byref(rd), stream)
        if rc != SUCCESS and (check or rc != 5):
            raise HipError(rc, "honei_hip_mg_solve_mixed", last_error())
        return o.value, i.value, rd.value

    def __del__(self):
        try:
            if self.h:
                lib().honei_hip_mg_destroy(self.h)
                self.h = None
        except Exception:
            pass


class BatchConfig(C.Structure):
    """honei_hip_batch_config (include/honei_hip.h)."""
    _fields_ = [("inner_vtype", C.c_int32), ("n_pre", C.c_int32), ("n_post", C.c_int32), ("omega", _d),
                ("tolerance", _d), ("max_outer", C.c_int32), ("averaging", C.c_int32),
                ("iters_per_launch", C.c_int32), ("autotune", C.c_int32), ("outer_method", C.c_int32),
                ("fused_levels", C.c_int32), ("use_graphs", C.c_int32), ("pair_sweeps", C.c_int32),
                ("matrix_free", C.c_int32), ("first_launch_from_history", C.c_int32)]


class SlotInfo(C.Structure):
    """honei_hip_slot_info (include/honei_hip.h)."""
    _fields_ = [("initial", _d), ("norm", _d), ("rel", _d), ("state", C.c_int32), ("iters", C.c_int32),
                ("status", C.c_int32), ("level", C.c_int32)]


def batch_config(**kw):
    c = BatchConfig()
    call("honei_hip_batch_config_default", C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise TypeError(f"unknown batch option {k}")
        setattr(c, k, v)
    return c


class Request(C.Structure):
    """honei_hip_request (include/honei_hip.h)."""
    _fields_ = [("arrival_s", _d), ("level", C.c_int32), ("pad", C.c_int32), ("coeff_seed", _u64),
                ("rhs_seed", _u64), ("x_dev", _p), ("coeff_host", _p), ("b_host", _p), ("x_host", _p)]


class RequestResult(C.Structure):
    """honei_hip_request_result (include/honei_hip.h)."""
    _fields_ = [("admit_s", _d), ("done_s", _d), ("rel", _d), ("iters", C.c_int32), ("status", C.c_int32),
                ("slot", C.c_int64)]


class ServeStats(C.Structure):
    """honei_hip_serve_stats (include/honei_hip.h)."""
    _fields_ = [("completed", C.c_int64), ("not_converged", C.c_int64), ("rejected", C.c_int64),
                ("launches", C.c_int64), ("max_queue", C.c_int64), ("wall_s", _d), ("solves_per_s", _d),
                ("p50_ms", _d), ("p99_ms", _d), ("max_ms", _d), ("mean_ms", _d), ("device_bytes", _u64)]


class ServeConfig(C.Structure):
    """honei_hip_serve_config (include/honei_hip.h)."""
    _fields_ = [("batch", BatchConfig), ("nclass", C.c_int32), ("max_slots_per_class", C.c_int32),
                ("levels", C.c_int64 * 16), ("counts", C.c_int64 * 16), ("budget_bytes", _u64), ("sigma", _d)]


class PinnedArray:
    """Pinned host memory (honei_hip_host_alloc) with a numpy view."""

    def __init__(self, n: int, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        p = _p()
        call("honei_hip_host_alloc", C.byref(p), max(1, self.n * self.dtype.itemsize))
        self.ptr = p.value
        buf = (C.c_char * max(1, self.n * self.dtype.itemsize)).from_address(self.ptr)
        self.a = np.frombuffer(buf, dtype=self.dtype, count=self.n)

    def free(self):
        if getattr(self, "ptr", None):
            self.a = None
            lib().honei_hip_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def serve(levels, requests, counts=None, budget_bytes=0, sigma=1.0, max_slots_per_class=256, **batch_kw):
    """Run the continuous-batching loop over `requests` (list of
    (arrival_s, level, coeff_seed, rhs_seed[, DeviceArray x[, PinnedArray
    coeff, PinnedArray b, PinnedArray x_host]])).  Returns (list of
    RequestResult, ServeStats)."""
    cfg = ServeConfig()
    cfg.batch = batch_config(**batch_kw)
    cfg.nclass = len(levels)
    cfg.max_slots_per_class = max_slots_per_class
    for i, lv in enumerate(levels):
        cfg.levels[i] = lv
        cfg.counts[i] = 0 if counts is None else counts[i]
    cfg.budget_bytes = budget_bytes
    cfg.sigma = sigma
    rq = (Request * max(1, len(requests)))()
    for i, r in enumerate(requests):
        rq[i].arrival_s, rq[i].level, rq[i].coeff_seed, rq[i].rhs_seed = r[0], r[1], r[2], r[3]
        rq[i].x_dev = r[4].ptr if len(r) > 4 and r[4] is not None else None
        if len(r) > 5:
            rq[i].coeff_host = r[5].ptr if r[5] is not None else None
            rq[i].b_host = r[6].ptr if r[6] is not None else None
            rq[i].x_host = r[7].ptr if len(r) > 7 and r[7] is not None else None
    res = (RequestResult * max(1, len(requests)))()
    st = ServeStats()
    call("honei_hip_serve_run", C.byref(cfg), rq, len(requests), res, C.byref(st))
    return list(res)[:len(requests)], st


def fill_uniform(arr: "DeviceArray", lo, hi, seed, stream=None):
    call("honei_hip_fill_uniform", arr.ptr, arr.n, lo, hi, seed, stream)


def fill_lognormal(arr: "DeviceArray", sigma, seed, stream=None):
    call("honei_hip_fill_lognormal", arr.ptr, arr.n, sigma, seed, stream)


class Batch:
    """Owning handle to a batched multigrid pool (honei_hip_batch_*)."""

    def __init__(self, levels, counts, cfg: BatchConfig = None, budget_bytes=0):
        cfg = cfg if cfg is not None else batch_config()
        lv = (C.c_int64 * len(levels))(*levels)
        ct = (C.c_int64 * len(counts))(*counts)
        p = _p()
        call("honei_hip_batch_create", C.byref(p), lv, ct, len(levels), C.byref(cfg), budget_bytes)
        self.h = p.value
        n, b = C.c_int64(), _u64()
        call("honei_hip_batch_info", self.h, C.byref(n), C.byref(b), None)
        self.nslots, self.device_bytes = n.value, b.value
        self.levels, self.counts = list(levels), list(counts)

    def slot(self, cls, j):
        g = C.c_int64()
        call("honei_hip_batch_slot", self.h, cls, j, C.byref(g))
        return g.value

    def load(self, slot, coeff: "DeviceArray", b: "DeviceArray"):
        call("honei_hip_batch_load", self.h, slot, coeff.ptr, b.ptr)

    def load_many(self, slots, coeffs, bs):
        n = len(slots)
        sl = (C.c_int64 * n)(*slots)
        cp = (_p * n)(*[c.ptr for c in coeffs])
        bp = (_p * n)(*[b.ptr for b in bs])
        call("honei_hip_batch_load_many", self.h, n, sl, cp, bp)

    def load_seeded(self, slots, coeff_seeds, rhs_seeds, sigma=1.0):
        n = len(slots)
        call("honei_hip_batch_load_seeded", self.h, n, (C.c_int64 * n)(*slots), (_u64 * n)(*coeff_seeds),
             (_u64 * n)(*rhs_seeds), sigma)

    def activate(self, wait=True):
        call("honei_hip_batch_activate", self.h, 1 if wait else 0)

    def step(self):
        call("honei_hip_batch_step", self.h)

    def poll(self):
        info = (SlotInfo * self.nslots)()
        call("honei_hip_batch_poll", self.h, info)
        return info

    def result(self, slot, x: "DeviceArray"):
        call("honei_hip_batch_result", self.h, slot, x.ptr)

    def run(self):
        k = C.c_int64()
        call("honei_hip_batch_run", self.h, C.byref(k))
        return k.value

    def shape(self, level):
        """(threads per row, block size, slots per thread) of an inner level."""
        t, b, g = C.c_int32(), C.c_int32(), C.c_int32()
        call("honei_hip_batch_shape", self.h, level, C.byref(t), C.byref(b), C.byref(g), None, None)
        return t.value, b.value, g.value

    def tuning(self, level):
        """(threads, block, slots per thread, measured sweep us, tuning-batch slots)."""
        t, b, g, us, ns = C.c_int32(), C.c_int32(), C.c_int32(), _d(), C.c_int64()
        call("honei_hip_batch_shape", self.h, level, C.byref(t), C.byref(b), C.byref(g), C.byref(us), C.byref(ns))
        return t.value, b.value, g.value, us.value, ns.value

    def pair_timing(self, level):
        """(paired Jacobi0 + defect us, paired two-sweep us) from the autotuner
        over the tuning batch of `tuning(level)`; zeros when not paired."""
        a, b = _d(), _d()
        call("honei_hip_batch_pair_timing", self.h, level, C.byref(a), C.byref(b))
        return a.value, b.value

    def time_level(self, level, reps=20):
        """(paired pre, restriction, prolongation, paired post, FP64 q = A p)
        microseconds per launch over every slot present on `level`
        (honei_hip_batch_time_level; overwrites the pool's vectors)."""
        us = (_d * 5)()
        call("honei_hip_batch_time_level", self.h, level, reps, us)
        return tuple(us)

    def launches(self):
        k = C.c_int64()
        call("honei_hip_batch_info", self.h, None, None, C.byref(k))
        return k.value

    def close(self):
        if getattr(self, "h", None):
            lib().honei_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cg_solve(D: DeviceEll, b: DeviceArray, x: DeviceArray, max_iters=10000, eps=1e-8, stream=None, check=True):
    used, fd = _sz(), _d()
    rc = lib().honei_hip_cg_solve(D.vtype, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, D.rows, D.ncpr, D.stride, D.threads,
                                  b.ptr, x.ptr, max_iters, eps, C.byref(used), C.byref(fd), stream)
    if rc != SUCCESS and (check or rc != 5):
        raise HipError(rc, "honei_hip_cg_solve", last_error())
    return used.value, fd.value


def cg_solve_noalloc(D: DeviceEll, b: DeviceArray, x: DeviceArray, max_iters=10000, eps=1e-8, stream=None,
                     check=True):
    """honei_hip_cg_solve_noalloc: cg_solve on the library's fixed device
    scratch (no allocation at call time; rows bounded by that scratch)."""
    used, fd = _sz(), _d()
    rc = lib().honei_hip_cg_solve_noalloc(D.vtype, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, D.rows, D.ncpr, D.stride,
                                          D.threads, b.ptr, x.ptr, max_iters, eps, C.byref(used), C.byref(fd),
                                          stream)
    if rc != SUCCESS and (check or rc != 5):
        raise HipError(rc, "honei_hip_cg_solve_noalloc", last_error())
    return used.value, fd.value


def pcg_jacobi(D: DeviceEll, b: DeviceArray, x: DeviceArray, dd_inverted: DeviceArray, max_iters=10000, eps=1e-8,
               stream=None, check=True):
    """ConjugateGradients<tags::GPU::HIP, JAC>::value(A, b, x, dd_inverted,
    max_iters, used_iters, eps_relative) (math/conjugate_gradients.hh:
    1110-1154).  Returns (used_iters, final ||r||); with check=False a run
    that reaches max_iters returns normally, as the reference does."""
    used, fd = _sz(), _d()
    rc = lib().honei_hip_pcg_jacobi(D.vtype, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, D.rows, D.ncpr, D.stride, D.threads,
                                    b.ptr, x.ptr, dd_inverted.ptr, max_iters, eps, C.byref(used), C.byref(fd),
                                    stream)
    if rc != SUCCESS and (check or rc != 5):
        raise HipError(rc, "honei_hip_pcg_jacobi", last_error())
    return used.value, fd.value
