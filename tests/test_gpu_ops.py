"""GPU parity of the ELL operators and BLAS-1/transfer kernels, through the C ABI.

Bar: bit-exact where the arithmetic order matches the CPU path (T = 1 ELL
products/defects/Jacobi, elementwise ops, geometric transfers); for T > 1
(tree reduction over lanes, as the reference's CUDA kernel) the error per
row is bounded by 1e-6 (FP32) / 1e-13 (FP64) relative to sum_j |a_rj x_j|.
Reductions accumulate in double and are compared to an exact (math.fsum)
sum with a 1e-12 relative bound.
"""
import math

import numpy as np
import pytest

import honei_amd as H
import oracle
from tests.patterns import honei_pattern_csr, random_band_csr, uniform_pm1

pytestmark = pytest.mark.gpu

TOL = {np.float32: 1e-6, np.float64: 1e-13}


@pytest.fixture(scope="module", autouse=True)
def _init():
    H.init(0)
    yield


def _abs_row_sums(A, x):
    """sum_j |a_rj x_j| per row, in float64."""
    Aa = oracle.Ell(A.rows, A.cols, A.stride, A.ncpr, A.threads, A.Aj, np.abs(A.Ax.astype(np.float64)), A.Arl)
    return oracle.product(Aa, np.abs(x.astype(np.float64)))


def _gpu_product(A, x, vtype=None, **kw):
    D = H.DeviceEll.from_oracle(A, vtype)
    dx = H.DeviceArray.from_host(x, D.vec_dtype)
    dy = H.DeviceArray(A.rows, D.vec_dtype)
    D.product(dy, dx, **kw)
    return dy.get()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("threads", [1, 2, 4, 8, 16, 32, 64])
def test_product_random_band(dtype, threads):
    n = 20000
    print(f"product random band n={n} T={threads} {np.dtype(dtype).name}", flush=True)
    rp, col, val = random_band_csr(n, seed=21, max_len=27, band=4096, dtype=dtype)
    A = oracle.ell_from_sm(n, n, rp, col, val, threads)
    x = uniform_pm1(2, n).astype(dtype)
    y = _gpu_product(A, x)
    yo = oracle.product(A, x)
    if threads == 1:
        assert np.array_equal(y.view(np.uint8), yo.view(np.uint8))
    else:
        err = np.abs(y.astype(np.float64) - yo.astype(np.float64))
        assert (err <= TOL[dtype] * _abs_row_sums(A, x) + 1e-30).all()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("side", [3, 33, 129, 513])
def test_product_q1_bitexact(dtype, side):
    print(f"product q1 side={side} {np.dtype(dtype).name}", flush=True)
    A = oracle.q1_matrix(side, dtype)
    x = uniform_pm1(1, side * side).astype(dtype)
    assert np.array_equal(_gpu_product(A, x), oracle.product(A, x))


@pytest.mark.parametrize("threads", [1, 4, 16])
@pytest.mark.parametrize("size", [11, 22, 88, 352])
def test_product_reference_pattern(size, threads):
    """la/product_TEST.cc:786-803 (x = 4, x0 = 1, x1 = 2; size x size+3)."""
    for dtype in (np.float32, np.float64):
        rp, col, val = honei_pattern_csr(size, size + 3, dtype=dtype)
        A = oracle.ell_from_sm(size, size + 3, rp, col, val, threads)
        x = np.full(size + 3, 4, dtype)
        x[0], x[1] = 1, 2
        y = _gpu_product(A, x)
        yo = oracle.product(A, x)
        err = np.abs(y.astype(np.float64) - yo)
        assert (err <= TOL[dtype] * _abs_row_sums(A, x) + 1e-30).all()


def _idx64_arrays(A):
    """Device copies of the matrix exactly as the reference's glue hands them
    to its launcher: Aj and Arl as the 64-bit unsigned long arrays of
    SparseMatrixELL (la/product-cuda.cc:178-194), Ax in its value type."""
    return (H.DeviceArray.from_host(A.Aj, np.uint64), H.DeviceArray.from_host(A.Ax, A.Ax.dtype),
            H.DeviceArray.from_host(A.Arl, np.uint64))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("threads", [1, 2, 4, 8, 16, 32, 64])
def test_product_and_defect_on_reference_index_arrays(dtype, threads):
    """honei_hip_{product,defect}_smell_dv_*_idx64 read the 64-bit Aj / Arl
    as the reference stores them: the same bits as the int32 entries (and
    the oracle's at T = 1), sub-ranges included; the narrowed arrays are the
    int32 ones."""
    n = 20000
    rp, col, val = random_band_csr(n, seed=31, max_len=27, band=4096, dtype=dtype)
    A = oracle.ell_from_sm(n, n, rp, col, val, threads)
    x = uniform_pm1(4, n).astype(dtype)
    rhs = uniform_pm1(5, n).astype(dtype)
    sfx = "float" if dtype == np.float32 else "double"
    Aj, Ax, Arl = _idx64_arrays(A)
    D = H.DeviceEll.from_oracle(A)
    dx, drhs = H.DeviceArray.from_host(x), H.DeviceArray.from_host(rhs)
    y32, y64 = H.DeviceArray(n, dtype), H.DeviceArray(n, dtype)
    D.product(y32, dx)
    H.call(f"honei_hip_product_smell_dv_{sfx}_idx64", dx.ptr, y64.ptr, Aj.ptr, Ax.ptr, Arl.ptr, 0, n, A.ncpr,
           A.stride, 0, threads, None)
    assert np.array_equal(y64.get().view(np.uint8), y32.get().view(np.uint8))
    if threads == 1:
        assert np.array_equal(y64.get().view(np.uint8), oracle.product(A, x).view(np.uint8))
    ys = H.DeviceArray(n, dtype)
    H.call(f"honei_hip_product_smell_dv_{sfx}_idx64", dx.ptr, ys.ptr, Aj.ptr, Ax.ptr, Arl.ptr, 777, 15000, A.ncpr,
           A.stride, 256, threads, None)
    assert np.array_equal(ys.get()[:15000 - 777], y32.get()[777:15000])
    r32, r64 = H.DeviceArray(n, dtype), H.DeviceArray(n, dtype)
    D.defect(r32, drhs, dx)
    H.call(f"honei_hip_defect_smell_dv_{sfx}_idx64", drhs.ptr, r64.ptr, Aj.ptr, Ax.ptr, Arl.ptr, dx.ptr, 0, n,
           A.ncpr, A.stride, 0, threads, None)
    assert np.array_equal(r64.get().view(np.uint8), r32.get().view(np.uint8))
    nj = H.DeviceArray(Aj.n, np.uint32)
    H.call("honei_hip_narrow_indices", nj.ptr, Aj.ptr, Aj.n, n, None)
    assert np.array_equal(nj.get(), D.Aj.get().view(np.uint32))
    with pytest.raises(H.HipError) as e:
        H.call("honei_hip_narrow_indices", nj.ptr, Aj.ptr, Aj.n, n // 2, None)
    assert e.value.code == 1


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_product_on_reference_index_arrays_q1_2049(dtype):
    """Full size (Q1 2049^2, the fixed-width buffer-load row sum): the 64-bit
    index entry is bit-exact against the CPU path."""
    A = oracle.q1_matrix(2049, dtype)
    x = uniform_pm1(2, A.rows).astype(dtype)
    sfx = "float" if dtype == np.float32 else "double"
    Aj, Ax, Arl = _idx64_arrays(A)
    dx, dy = H.DeviceArray.from_host(x), H.DeviceArray(A.rows, dtype)
    H.call(f"honei_hip_product_smell_dv_{sfx}_idx64", dx.ptr, dy.ptr, Aj.ptr, Ax.ptr, Arl.ptr, 0, A.rows, A.ncpr,
           A.stride, 0, 1, None)
    assert np.array_equal(dy.get().view(np.uint8), oracle.product(A, x).view(np.uint8))


def sell_padded_slots(lengths, window=2048, c=64, keys=64):
    """Slots of the sliced copy: rows sorted by min(length, keys - 1)
    descending, ties by row, within windows; each slice of c sorted rows
    padded to its longest row."""
    lengths = np.asarray(lengths, np.int64)
    key = np.minimum(lengths, keys - 1)
    n = lengths.size
    perm = np.concatenate([np.arange(w0, min(n, w0 + window))[np.lexsort(
        (np.arange(w0, min(n, w0 + window)), -key[w0:min(n, w0 + window)]))] for w0 in range(0, n, window)])
    L = lengths[perm]
    return int(sum(c * L[s:s + c].max() for s in range(0, n, c)))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,max_len", [(1, 27), (63, 27), (2048, 27), (2049, 27), (20000, 27), (5000, 100),
                                       (300001, 27)])
def test_sliced_copy_product_bitexact(dtype, n, max_len):
    """honei_hip_product_sell on the sliced copy of a ragged random ELL
    matrix: y bit-identical to the ELL product and to Product<tags::CPU>
    (every row's terms in stored order), with x staged in LDS per window and
    gathered from global memory; the padded slot count is the one the
    windowed sort implies (rows longer than 63 share the last key).  Bands
    up to 300,000 columns: windows wider than the LDS stage fall back to
    global gathers."""
    band = max(n - 1, 1) if n == 300001 and dtype == np.float64 else min(4096, max(n - 1, 1))
    rp, col, val = random_band_csr(n, seed=40 + n % 7, max_len=max_len, band=band, dtype=dtype)
    A = oracle.ell_from_sm(n, n, rp, col, val, 1)
    x = uniform_pm1(6, n).astype(dtype)
    D = H.DeviceEll.from_oracle(A)
    S = H.Sell(D)
    dx = H.DeviceArray.from_host(x)
    ye, ys = H.DeviceArray(n, dtype), H.DeviceArray(n, dtype)
    D.product(ye, dx)
    S.product(ys, dx)
    assert np.array_equal(ys.get().view(np.uint8), ye.get().view(np.uint8))
    assert np.array_equal(ys.get().view(np.uint8), oracle.product(A, x).view(np.uint8))
    # x gathered from global memory instead of each window's LDS copy: the same bits
    yg = H.DeviceArray(n, dtype)
    S.stage_x(False)
    S.product(yg, dx)
    assert np.array_equal(yg.get().view(np.uint8), ye.get().view(np.uint8))
    assert S.padded_slots == sell_padded_slots(A.Arl)
    # no more than the padded ELL, rounded up to whole 64-row slices
    assert S.padded_slots <= A.ncpr * max(A.stride, 64 * ((n + 63) // 64))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sliced_copy_of_the_random_4m_matrix(dtype):
    """The SpMV benchmark's random matrix (4,194,304 rows, lengths U{1..27},
    columns within +-4096, generated in HBM): the sliced product equals the
    ELL product bit for bit, and the sliced copy stores under 55 % of the
    padded ELL slots."""
    n, ncpr = 1 << 22, 27
    vt = H.F64 if dtype == np.float64 else H.F32
    Aj = H.DeviceArray(ncpr * n, np.int32)
    Ax = H.DeviceArray(ncpr * n, dtype)
    Arl = H.DeviceArray(n, np.uint32)
    H.call("honei_hip_random_ell", vt, n, 27, 4096, 2, Aj.ptr, Ax.ptr, Arl.ptr, n, 1, None)
    D = H.DeviceEll.__new__(H.DeviceEll)
    D.rows, D.cols, D.stride, D.ncpr, D.threads, D.vtype = n, n, n, ncpr, 1, vt
    D.Aj, D.Ax, D.Arl = Aj, Ax, Arl
    S = H.Sell(D)
    x = H.DeviceArray.from_host(uniform_pm1(2, n).astype(dtype))
    ye, ys = H.DeviceArray(n, dtype), H.DeviceArray(n, dtype)
    D.product(ye, x)
    S.product(ys, x)
    assert np.array_equal(ys.get().view(np.uint8), ye.get().view(np.uint8))
    S.stage_x(False)
    S.product(ys, x)
    assert np.array_equal(ys.get().view(np.uint8), ye.get().view(np.uint8))
    assert S.padded_slots == sell_padded_slots(Arl.get())
    assert S.padded_slots < 0.55 * ncpr * n


def test_product_row_range_offsets_output():
    """Sub-range convention of cuda_product_smell_dv_*: y[r - row_start]."""
    A = oracle.q1_matrix(65)
    x = uniform_pm1(3, A.rows)
    D = H.DeviceEll.from_oracle(A)
    dx = H.DeviceArray.from_host(x)
    dy = H.DeviceArray(1000, np.float64)
    D.product(dy, dx, row_start=1500, row_end=2500)
    assert np.array_equal(dy.get(), oracle.product(A, x)[1500:2500])


def test_product_empty_and_single_row():
    A = oracle.q1_matrix(3)
    D = H.DeviceEll.from_oracle(A)
    dx = H.DeviceArray.from_host(np.ones(9))
    dy = H.DeviceArray(9, np.float64)
    dy.upload(np.full(9, 7.0))
    D.product(dy, dx, row_start=4, row_end=4)   # empty range: no write
    assert (dy.get() == 7.0).all()
    D.product(dy, dx, row_start=4, row_end=5)
    assert dy.get()[0] == pytest.approx(8 / 3 - 8 / 3, abs=1e-15)


def test_product_bf16_values():
    """BF16 storage of matrix values, FP32 accumulation: equal to the FP32
    product of the bf16-rounded matrix."""
    n = 30000
    rp, col, val = random_band_csr(n, seed=5, max_len=27, band=2000, dtype=np.float32)
    A = oracle.ell_from_sm(n, n, rp, col, val, 1)
    x = uniform_pm1(6, n).astype(np.float32)
    y = _gpu_product(A, x, vtype=H.BF16)
    u = A.Ax.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    Ab = oracle.Ell(A.rows, A.cols, A.stride, A.ncpr, 1, A.Aj, r.astype(np.uint32).view(np.float32), A.Arl)
    assert np.array_equal(y, oracle.product(Ab, x))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("threads", [1, 8])
def test_defect_and_fused_norm(dtype, threads):
    n = 50000
    rp, col, val = random_band_csr(n, seed=8, max_len=12, band=300, dtype=dtype)
    A = oracle.ell_from_sm(n, n, rp, col, val, threads)
    x = uniform_pm1(1, n).astype(dtype)
    b = uniform_pm1(2, n).astype(dtype)
    D = H.DeviceEll.from_oracle(A)
    dx, db = H.DeviceArray.from_host(x), H.DeviceArray.from_host(b)
    dr = H.DeviceArray(n, dtype)
    D.defect(dr, db, dx)
    r = dr.get()
    ro = oracle.defect(A, b, x)
    if threads == 1:
        assert np.array_equal(r, ro)
    else:
        assert (np.abs(r.astype(np.float64) - ro) <= TOL[dtype] * (_abs_row_sums(A, x) + np.abs(b)) + 1e-30).all()
    dn = H.DeviceArray(1, np.float64)
    dr2 = H.DeviceArray(n, dtype)
    D.defect_norm(dr2, db, dx, dn)
    assert np.array_equal(dr2.get(), r)
    exact = math.fsum((r.astype(np.float64) ** 2).tolist())
    assert dn.get()[0] == pytest.approx(exact, rel=1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_jacobi_sweeps_match_oracle(dtype):
    """Fused sweep == Defect then ScaledSum(x, t, dinv) (math/jacobi.hh:268-279);
    JACOBI0 == ElementProduct first sweep followed by one sweep."""
    side = 257
    A = oracle.q1_matrix(side, dtype)
    n = side * side
    rhs = uniform_pm1(4, n).astype(dtype)
    dinv = (dtype(0.7) / oracle.q1_bands(side, dtype)[4]).astype(dtype)
    x = uniform_pm1(5, n).astype(dtype)
    D = H.DeviceEll.from_oracle(A)
    dx, drhs, ddinv = H.DeviceArray.from_host(x), H.DeviceArray.from_host(rhs), H.DeviceArray.from_host(dinv)
    dxn = H.DeviceArray(n, dtype)
    D.jacobi(dxn, dx, drhs, ddinv)
    expect = oracle.scaled_sum_v(x, oracle.defect(A, rhs, x), dinv)
    assert np.array_equal(dxn.get(), expect)
    # from zero, two sweeps fused, accumulated into acc
    acc0 = uniform_pm1(6, n).astype(dtype)
    dacc = H.DeviceArray.from_host(acc0)
    D.jacobi(dxn, None, drhs, ddinv, acc=dacc)
    c1 = (dinv * rhs).astype(dtype)
    c2 = oracle.scaled_sum_v(c1, oracle.defect(A, rhs, c1), dinv)
    assert np.array_equal(dxn.get(), c2)
    assert np.array_equal(dacc.get(), (acc0 + c2).astype(dtype))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [0, 1, 63, 64, 1000, 1 << 20, 3_000_001])
def test_reductions(dtype, n):
    vt = H.F32 if dtype == np.float32 else H.F64
    x = uniform_pm1(10, n).astype(dtype)
    y = uniform_pm1(11, n).astype(dtype)
    dx, dy = H.DeviceArray.from_host(x), H.DeviceArray.from_host(y)
    out = H.DeviceArray(1, np.float64)
    H.call("honei_hip_dot_product_dev", vt, out.ptr, dx.ptr, dy.ptr, n, None)
    ex = math.fsum((x.astype(np.float64) * y.astype(np.float64)).tolist())
    assert out.get()[0] == pytest.approx(ex, rel=1e-12, abs=1e-300)
    H.call("honei_hip_norm_l2_dev", vt, out.ptr, dx.ptr, n, None)
    ex2 = math.fsum((x.astype(np.float64) ** 2).tolist())
    assert out.get()[0] == pytest.approx(ex2, rel=1e-12, abs=1e-300)
    # repeated calls reuse the arrival counter: results stay identical
    v1 = out.get()[0]
    for _ in range(3):
        H.call("honei_hip_norm_l2_dev", vt, out.ptr, dx.ptr, n, None)
        assert out.get()[0] == v1


def test_host_reductions_and_axpy_norm():
    n = 100003
    x = uniform_pm1(1, n)
    y = uniform_pm1(2, n)
    dx, dy = H.DeviceArray.from_host(x), H.DeviceArray.from_host(y)
    import ctypes as C
    r = C.c_double()
    H.call("honei_hip_dot_product_two_double", C.byref(r), dx.ptr, dy.ptr, n, None)
    assert r.value == pytest.approx(math.fsum((x * y).tolist()), rel=1e-12)
    rf = C.c_float()
    dxf = H.DeviceArray.from_host(x.astype(np.float32))
    H.call("honei_hip_norm_l2_one_float", C.byref(rf), dxf.ptr, n, None)
    assert rf.value == pytest.approx(float(np.sum(x.astype(np.float32).astype(np.float64) ** 2)), rel=1e-6)
    out = H.DeviceArray(1, np.float64)
    H.call("honei_hip_axpy_norm_dev", H.F64, dx.ptr, dy.ptr, -0.25, out.ptr, n, None)
    xn = x + y * -0.25
    assert np.array_equal(dx.get(), xn)
    assert out.get()[0] == pytest.approx(math.fsum((xn * xn).tolist()), rel=1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_elementwise_bitexact(dtype):
    vt = H.F32 if dtype == np.float32 else H.F64
    n = 777777
    x = uniform_pm1(1, n).astype(dtype)
    y = uniform_pm1(2, n).astype(dtype)
    z = (uniform_pm1(3, n) + 2).astype(dtype)
    s = dtype(0.37)

    def run(name, *arrays_scalar, out=0):
        devs = [H.DeviceArray.from_host(a) if isinstance(a, np.ndarray) else a for a in arrays_scalar]
        args = [d.ptr if isinstance(d, H.DeviceArray) else d for d in devs]
        H.call(name, vt, *args, n, None)
        return devs[out].get()

    assert np.array_equal(run("honei_hip_scaled_sum_two", x.copy(), y, float(s)), x + y * s)
    assert np.array_equal(run("honei_hip_scaled_sum_three_s", np.zeros_like(x), x, y, float(s)), x + y * s)
    assert np.array_equal(run("honei_hip_scaled_sum_three", x.copy(), y, z), x + y * z)
    assert np.array_equal(run("honei_hip_sum_two", x.copy(), y), x + y)
    assert np.array_equal(run("honei_hip_difference_two", x.copy(), y), x - y)
    assert np.array_equal(run("honei_hip_difference_three", np.zeros_like(x), x, y), x - y)
    assert np.array_equal(run("honei_hip_scale_one", x.copy(), float(s)), x * s)
    assert np.array_equal(run("honei_hip_element_product_three", np.zeros_like(x), x, y), x * y)
    assert np.array_equal(run("honei_hip_element_inverse_one", z.copy()), (dtype(1) / z).astype(dtype))


def test_convert_roundtrip():
    n = 123457
    x = uniform_pm1(9, n) * 1e3
    dx = H.DeviceArray.from_host(x)
    df = H.DeviceArray(n, np.float32)
    dd = H.DeviceArray(n, np.float64)
    H.call("honei_hip_convert_double_float", df.ptr, dx.ptr, n, None)
    H.call("honei_hip_convert_float_double", dd.ptr, df.ptr, n, None)
    assert np.array_equal(df.get(), x.astype(np.float32))
    assert np.array_equal(dd.get(), x.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("level", [1, 2, 3, 5, 8, 10])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_transfers_bitexact(level, dtype):
    vt = H.F32 if dtype == np.float32 else H.F64
    sc, sf = 2 ** level + 1, 2 ** (level + 1) + 1
    c = uniform_pm1(level, sc * sc).astype(dtype)
    f = uniform_pm1(level + 50, sf * sf).astype(dtype)
    for mask in [(2,) * 8, (2, 2, 2, 2, 2, 1, 2, 2), (1,) * 8, (2, 1, 2, 1, 1, 2, 1, 2)]:
        m = np.asarray(mask, np.uint64)
        dc, df = H.DeviceArray.from_host(c), H.DeviceArray(sf * sf, dtype)
        H.call("honei_hip_prolongation", vt, df.ptr, sf * sf, dc.ptr, sc * sc, m.ctypes.data, None)
        assert np.array_equal(df.get(), oracle.prolongation(c, mask))
        df2, dc2 = H.DeviceArray.from_host(f), H.DeviceArray(sc * sc, dtype)
        H.call("honei_hip_restriction", vt, dc2.ptr, sc * sc, df2.ptr, sf * sf, m.ctypes.data, None)
        assert np.array_equal(dc2.get(), oracle.restriction(f, mask))


def test_invalid_arguments_are_rejected():
    A = oracle.q1_matrix(5)
    D = H.DeviceEll.from_oracle(A)
    dx = H.DeviceArray.from_host(np.ones(25))
    dy = H.DeviceArray(25, np.float64)
    with pytest.raises(H.HipError) as e:
        H.call("honei_hip_product_smell_dv_double", dx.ptr, dy.ptr, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, 0, 25,
               D.ncpr, D.stride, 0, 3, None)
    assert e.value.code == 1
    with pytest.raises(H.HipError):
        H.call("honei_hip_product_smell_dv_double", dx.ptr, dy.ptr, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, 0, 25,
               D.ncpr, 8, 0, 1, None)    # stride < rows
    m = np.full(8, 2, np.uint64)
    with pytest.raises(H.HipError) as e:
        H.call("honei_hip_restriction", H.F64, dy.ptr, 24, dx.ptr, 25, m.ctypes.data, None)
    assert e.value.code == 2
    bad = np.array([0, 1, 99], np.uint64)
    buf = H.DeviceArray(3, np.int32)
    with pytest.raises(H.HipError):
        H.call("honei_hip_upload_indices", buf.ptr, bad.ctypes.data, 3, 25)


def _device_random_ell(n, vt, seed=2, max_len=27, band=4096, threads=1):
    stride = 32 * ((n * threads + 31) // 32)
    ncpr = (max_len + threads - 1) // threads
    Aj = H.DeviceArray(ncpr * stride, np.int32)
    Ax = H.DeviceArray(ncpr * stride, np.float64 if vt == H.F64 else np.float32)
    Arl = H.DeviceArray(n, np.uint32)
    H.call("honei_hip_random_ell", vt, n, max_len, band, seed, Aj.ptr, Ax.ptr, Arl.ptr, stride, threads, None)
    return oracle.Ell(n, n, stride, ncpr, threads, Aj.get().astype(np.uint64), Ax.get(),
                      Arl.get().astype(np.uint64))


@pytest.mark.parametrize("threads", [2, 4, 8, 32])
def test_random_ell_threads_layout_matches_conversion(threads):
    """The generator's T-interleaved layout equals SparseMatrixELL(T) of the
    same rows (CPU conversion of the T = 1 matrix)."""
    n = 3000
    A1 = _device_random_ell(n, H.F64, band=500)
    AT = _device_random_ell(n, H.F64, band=500, threads=threads)
    rp, col, val = A1.to_csr()
    C = oracle.ell_from_sm(n, n, rp, col, val, threads)
    assert (C.stride, C.ncpr) == (AT.stride, AT.ncpr)
    assert np.array_equal(C.Aj, AT.Aj) and np.array_equal(C.Ax, AT.Ax) and np.array_equal(C.Arl, AT.Arl)
    x = uniform_pm1(1, n)
    y = _gpu_product(AT, x)
    err = np.abs(y - oracle.product(AT, x))
    assert (err <= 1e-13 * _abs_row_sums(AT, x) + 1e-30).all()


@pytest.mark.parametrize("n,band", [(1000, 4096), (100000, 4096), (1 << 22, 4096)])
def test_random_ell_generator_structure(n, band):
    """Rows U{1..27} long, strictly increasing columns inside the clipped
    window, zero padding, non-zero values; same matrix for the same seed."""
    print(f"random ELL n={n}", flush=True)
    A = _device_random_ell(n, H.F64, band=band)
    L = A.Arl.astype(np.int64)
    assert L.min() >= 1 and L.max() <= 27
    M = A.Aj.reshape(27, A.stride)[:, :n].T.astype(np.int64)
    V = A.Ax.reshape(27, A.stride)[:, :n].T
    m = np.arange(27)[None, :] < L[:, None]
    assert ((np.diff(np.where(m, M, 1 << 40), axis=1) > 0) | ~m[:, 1:]).all()
    r = np.arange(n)[:, None]
    assert (np.where(m, np.abs(M - r), 0) <= band).all() and (M < n).all()
    assert (np.where(m, V != 0, True)).all() and (np.where(~m, (V == 0) & (M == 0), True)).all()
    if n >= 100000:
        assert abs(L.mean() - 14.0) < 0.05 and abs(V[m].std() - 1.0) < 0.01
    B = _device_random_ell(n, H.F64, band=band)
    assert np.array_equal(A.Aj, B.Aj) and np.array_equal(A.Ax, B.Ax)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_full_size_spmv_configs(dtype):
    """SpMV configurations at full size: Q1 2049^2 (4,198,401 rows) and the
    random 4,194,304-row ELL; T = 1 results bit-identical to the CPU path."""
    vt = H.F32 if dtype == np.float32 else H.F64
    print(f"full-size spmv {np.dtype(dtype).name}", flush=True)
    A = oracle.q1_matrix(2049, dtype)
    x = uniform_pm1(2, A.rows).astype(dtype)
    assert np.array_equal(_gpu_product(A, x), oracle.product(A, x))
    R = _device_random_ell(1 << 22, vt)
    xr = uniform_pm1(2, R.rows).astype(dtype)
    yr = oracle.product(R, xr)
    assert np.array_equal(_gpu_product(R, xr), yr)
    ys, _ = _staged(H.DeviceEll.from_oracle(R), xr, 1024, chunks=2)
    assert np.array_equal(ys, yr)


# ---- ELL product with x staged in LDS ---------------------------------------------------

def _staged(D, x, block, window=None, rpb=None, chunks=1):
    ranges, w = D.block_ranges(rpb or block // D.threads * chunks)
    dx = H.DeviceArray.from_host(x, D.vec_dtype)
    dy = H.DeviceArray(D.rows, D.vec_dtype)
    vs = 8 if D.vtype == H.F64 else 4
    D.product_staged(dy, dx, ranges, window or min(max(w, 1), 98304 // vs), blocksize=block, chunks=chunks)
    return dy.get(), w


@pytest.mark.parametrize("vtype", [H.F32, H.F64, H.BF16])
@pytest.mark.parametrize("threads", [1, 2, 4, 16])
@pytest.mark.parametrize("block,chunks", [(256, 1), (1024, 1), (512, 3)])
def test_staged_product_bitexact(vtype, threads, block, chunks):
    """Staged gathers give the bits of the plain kernel (and, for T = 1, of
    Product<tags::CPU>)."""
    n = 30000
    dtype = np.float64 if vtype == H.F64 else np.float32
    print(f"staged product n={n} T={threads} block={block} chunks={chunks} vt={vtype}", flush=True)
    rp, col, val = random_band_csr(n, seed=23, max_len=27, band=4096, dtype=dtype)
    A = oracle.ell_from_sm(n, n, rp, col, val, threads)
    D = H.DeviceEll.from_oracle(A, vtype)
    x = uniform_pm1(3, n).astype(dtype)
    ys, w = _staged(D, x, block, chunks=chunks)
    assert w <= block // threads * chunks + 2 * 4096
    yp = _gpu_product(A, x, vtype, blocksize=block)
    assert np.array_equal(ys.view(np.uint8), yp.view(np.uint8))
    if threads == 1 and vtype != H.BF16:
        assert np.array_equal(ys.view(np.uint8), oracle.product(A, x).view(np.uint8))


def test_staged_block_ranges_and_fallbacks():
    """Windows equal the per-block min/max columns; windows wider than the
    LDS allotment and ranges made for another block size fall back to global
    gathers with the same bits."""
    n = 5000
    rp, col, val = random_band_csr(n, seed=5, max_len=27, band=2000, dtype=np.float64)
    A = oracle.ell_from_sm(n, n, rp, col, val, 1)
    D = H.DeviceEll.from_oracle(A)
    ranges, w = D.block_ranges(256)
    got = ranges.get().reshape(-1, 2)
    for b in range(got.shape[0]):
        cols = col[rp[b * 256]:rp[min(n, (b + 1) * 256)]]
        cols = cols[cols != 0]   # column 0 stays out of the windows
        assert got[b, 0] == cols.min() and got[b, 1] == cols.max()
    assert w == int((got[:, 1] - got[:, 0] + 1).max())
    x = uniform_pm1(4, n)
    ref = oracle.product(A, x)
    for window, rpb in ((64, None), (w, 64)):
        y, _ = _staged(D, x, 256, window=window, rpb=rpb)
        assert np.array_equal(y.view(np.uint8), ref.view(np.uint8)), (window, rpb)
    dx, dy = H.DeviceArray.from_host(x), H.DeviceArray(n, np.float64)
    with pytest.raises(H.HipError):   # more LDS than allowed
        D.product_staged(dy, dx, ranges, 98304 // 8 + 1)
    coarse, _ = D.block_ranges(1024)
    with pytest.raises(H.HipError):   # ranges made for larger blocks: too few
        D.product_staged(dy, dx, coarse, w, blocksize=256)
    with pytest.raises(H.HipError):
        D.product_staged(dy, dx, coarse, w, blocksize=256, chunks=0)


# ---- banded Q1 product / defect --------------------------------------------------------

@pytest.mark.parametrize("side", [3, 33, 1025, 2049])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_banded_q1_bitexact(side, dtype):
    """Product<tags::CPU>(BandedMatrixQx) order reproduced: bit-identical to
    the oracle (itself bit-identical to the reference build); lognormal
    coefficient bands at the smaller sizes, the Poisson bands at all."""
    from tests.patterns import lognormal_cells
    n = side * side
    sfx = "float" if dtype == np.float32 else "double"
    x = uniform_pm1(side, n).astype(dtype)
    rhs = uniform_pm1(side + 1, n).astype(dtype)
    cases = [oracle.q1_bands(side, dtype)]
    if side <= 33:
        cases.append(oracle.q1_coeff_bands(side, lognormal_cells(side, (side - 1) ** 2), dtype))
    for bands in cases:
        db = [H.DeviceArray.from_host(np.ascontiguousarray(bands[k]), dtype) for k in range(9)]
        dx, dr = H.DeviceArray.from_host(x, dtype), H.DeviceArray.from_host(rhs, dtype)
        dy = H.DeviceArray(n, dtype)
        H.call(f"honei_hip_product_bmdv_q1_{sfx}", *[b.ptr for b in db], dx.ptr, dy.ptr, n, 0, side, None)
        assert np.array_equal(dy.get().view(np.uint8), oracle.banded_q1(bands, x).view(np.uint8))
        H.call(f"honei_hip_defect_q1_{sfx}", dr.ptr, *[b.ptr for b in db], dx.ptr, dy.ptr, n, 128, side, None)
        assert np.array_equal(dy.get().view(np.uint8), oracle.banded_q1(bands, x, rhs).view(np.uint8))


def test_banded_q1_invalid_arguments():
    n = 25
    d = [H.DeviceArray(n, np.float32) for _ in range(11)]
    with pytest.raises(H.HipError) as e:
        H.call("honei_hip_product_bmdv_q1_float", *[a.ptr for a in d], 24, 0, 5, None)
    assert e.value.code == 2
    with pytest.raises(H.HipError) as e:
        H.call("honei_hip_product_bmdv_q1_float", *[a.ptr for a in d[:9]], d[9].ptr, d[10].ptr, n, 96, 5, None)
    assert e.value.code == 1
