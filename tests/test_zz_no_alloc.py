"""honei_hip_cg_solve_noalloc against honei_hip_cg_solve.

The allocation-free entry point runs the same kernels with the same launch
shapes on a fixed library-held scratch array, so its iterate, iteration
count and final defect are bit-identical to the allocating entry's on the
same inputs.  Two calls in a row on one device (the second one smaller,
taking the three-kernel iteration where the first took the five-kernel one)
check that nothing of the previous call's scratch leaks into the next.  The
FP64 true residual of each result is recomputed by the oracle.
"""
import numpy as np
import pytest

import honei_amd as H
import oracle
from tests.patterns import uniform_pm1

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    H.init(0)
    yield


def _system(side, seed, dtype):
    A = oracle.q1_matrix(side, np.float64)
    xt = uniform_pm1(seed, A.rows)
    b = oracle.product(A, xt)
    return A.astype(dtype), b.astype(dtype)


def _solve(fn, A, b, eps):
    D = H.DeviceEll.from_oracle(A)
    db = H.DeviceArray.from_host(b)
    dx = H.DeviceArray.from_host(np.zeros_like(b))
    used, fd = fn(D, db, dx, 5000, eps)
    return dx.get(), used, fd


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_equals_cg_solve_two_calls(dtype):
    eps = 1e-8 if dtype == np.float64 else 1e-5
    # 257^2 rows take the five-kernel iteration, 65^2 the three-kernel one
    for side, seed in ((257, 11), (65, 12)):
        A, b = _system(side, seed, dtype)
        x_ref, used_ref, fd_ref = _solve(H.cg_solve, A, b, eps)
        x_na, used_na, fd_na = _solve(H.cg_solve_noalloc, A, b, eps)
        assert used_na == used_ref, (side, used_na, used_ref)
        assert fd_na == fd_ref, (side, fd_na, fd_ref)
        assert np.array_equal(x_na.view(np.uint8), x_ref.view(np.uint8)), side
        A64, b64 = A.astype(np.float64), b.astype(np.float64)
        r = np.linalg.norm(oracle.defect(A64, b64, x_na.astype(np.float64)))
        assert r < 1.5 * eps * np.linalg.norm(b64), (side, r)


def test_noalloc_rejects_beyond_scratch():
    """Rows whose three vectors exceed the fixed scratch are refused before
    anything runs (the buffers are valid, zero-filled, in case one ran)."""
    rows = 2_000_000
    Aj = H.DeviceArray(rows, np.int32)
    Ax = H.DeviceArray(rows, np.float64)
    Arl = H.DeviceArray(rows, np.uint32)
    b = H.DeviceArray(rows, np.float64)
    x = H.DeviceArray(rows, np.float64)
    for a in (Aj, Ax, Arl, b, x):
        a.zero()
    used, fd = H._sz(), H._d()
    rc = H.lib().honei_hip_cg_solve_noalloc(H.F64, Aj.ptr, Ax.ptr, Arl.ptr, rows, 1, rows, 1, b.ptr, x.ptr, 10, 1e-8,
                                            H.C.byref(used), H.C.byref(fd), None)
    assert rc == 1
    assert "scratch" in H.last_error()
