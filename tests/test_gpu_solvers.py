"""GPU parity of the device Q1 assembly and the CG / multigrid solvers.

Bars:
  * Q1 assembly: index arrays and values bit-identical to the CPU path
    (FillMatrix -> SparseMatrix -> SparseMatrixELL).
  * Solvers: the GPU reductions accumulate in double in a tree while the CPU
    path sums sequentially in the working precision, so iterates agree to
    rounding, not bits.  Checked: the same iteration counts as the oracle
    (reference control flow), FP64 residual below the requested relative
    tolerance, and the solution within 1e-8 (FP64) / 1e-6 (mixed) relative
    of the oracle's; final error against the manufactured solution x_true
    is asserted as well.
"""
import os

import numpy as np
import pytest

import honei_amd as H
import oracle
from tests.patterns import uniform_pm1
from tests.spd import diag_inverse, lognormal_q1_spd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    H.init(0)
    yield


def _device_q1(side, vt, coeff=None, omega=0.7, threads=1):
    n = side * side
    stride = 32 * ((n * threads + 31) // 32)
    ncpr = (9 + threads - 1) // threads
    Aj = H.DeviceArray(ncpr * stride, np.int32)
    Ax = H.DeviceArray(ncpr * stride, {H.F32: np.float32, H.F64: np.float64, H.BF16: np.uint16}[vt])
    Arl = H.DeviceArray(n, np.uint32)
    dinv = H.DeviceArray(n, np.float64 if vt == H.F64 else np.float32)
    dc = H.DeviceArray.from_host(coeff, np.float64) if coeff is not None else None
    H.call("honei_hip_q1_assemble", vt, side, Aj.ptr, Ax.ptr, Arl.ptr, stride, None if dc is None else dc.ptr,
           dinv.ptr, omega, threads, None)
    return Aj.get(), Ax.get(), Arl.get(), dinv.get(), stride


@pytest.mark.parametrize("side", [3, 5, 33, 129, 1025])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_q1_assembly_bitexact(side, dtype):
    vt = H.F32 if dtype == np.float32 else H.F64
    Aj, Ax, Arl, dinv, stride = _device_q1(side, vt)
    A = oracle.q1_matrix(side, dtype)
    assert stride == A.stride and A.ncpr == (9 if side > 2 else 1)
    assert np.array_equal(Aj.astype(np.uint64), A.Aj)
    assert np.array_equal(Ax.view(np.uint8), A.Ax.view(np.uint8))
    assert np.array_equal(Arl.astype(np.uint64), A.Arl)
    d = oracle.q1_bands(side, dtype)[4]
    expect = ((np.ones(side * side, dtype) / d).astype(dtype) * dtype(0.7)).astype(dtype)
    assert np.array_equal(dinv, expect)


@pytest.mark.parametrize("side", [3, 17, 65])
@pytest.mark.parametrize("threads", [2, 4, 8])
def test_q1_assembly_threads_layout_bitexact(side, threads):
    """T-interleaved layout: the same arrays as SparseMatrixELL(SparseMatrix, T)."""
    Aj, Ax, Arl, _, stride = _device_q1(side, H.F32, threads=threads)
    A = oracle.q1_matrix(side, np.float32, threads=threads)
    assert stride == A.stride and len(Aj) == A.ncpr * A.stride
    assert np.array_equal(Aj.astype(np.uint64), A.Aj)
    assert np.array_equal(Ax.view(np.uint8), A.Ax.view(np.uint8))
    assert np.array_equal(Arl.astype(np.uint64), A.Arl)


def _numpy_variable_q1(side, kappa):
    """Brute-force assembly from element matrices (bilinear Q1, unit cells)."""
    n = side * side
    K = np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]]) / 6.0
    A = np.zeros((n, n))
    for ci in range(side - 1):
        for cj in range(side - 1):
            nodes = [ci * side + cj, ci * side + cj + 1, (ci + 1) * side + cj + 1, (ci + 1) * side + cj]
            A[np.ix_(nodes, nodes)] += kappa[ci * (side - 1) + cj] * K
    for r in range(n):
        i, j = divmod(r, side)
        if i in (0, side - 1) or j in (0, side - 1):
            A[r] = 0
            A[r, r] = 1
    return A


@pytest.mark.parametrize("side", [3, 5, 9, 17])
def test_q1_variable_coefficient_assembly(side):
    """No reference counterpart: pinned by brute-force element assembly and by
    kappa = 1 reproducing the reference's constant operator bit-for-bit."""
    kappa = np.exp(uniform_pm1(side, (side - 1) ** 2))
    Aj, Ax, Arl, dinv, stride = _device_q1(side, H.F64, coeff=kappa)
    E = oracle.Ell(side * side, side * side, stride, 9, 1, Aj.astype(np.uint64), Ax, Arl.astype(np.uint64))
    np.testing.assert_allclose(E.to_dense(), _numpy_variable_q1(side, kappa), rtol=1e-14, atol=1e-15)
    Aj1, Ax1, Arl1, _, _ = _device_q1(side, H.F64, coeff=np.ones((side - 1) ** 2))
    A = oracle.q1_matrix(side)
    assert np.array_equal(Ax1, A.Ax) and np.array_equal(Aj1.astype(np.uint64), A.Aj)


def _problem(side, seed):
    A = oracle.q1_matrix(side)
    xt = uniform_pm1(seed, side * side)
    return A, xt, oracle.product(A, xt)


# Q1 256^2 = 65,536 rows is the largest system whose CG iteration runs as
# three kernels (at most 256 SpMV blocks of 256 rows); 257^2 and 513^2 take
# the five-kernel iteration (one-block scalar kernels between the passes).
@pytest.mark.parametrize("dtype,side", [(np.float64, 33), (np.float64, 129), (np.float32, 33), (np.float32, 129),
                                        (np.float64, 256), (np.float64, 257), (np.float64, 513)])
def test_cg_matches_oracle(dtype, side):
    """CGSolver (math/cg.hh:63-133) on Q1 from x0 = 0: the oracle's iteration
    count (+-1 FP64, +-3 FP32), the FP64 true residual below the tolerance
    and the iterate within rounding of the oracle's."""
    A, xt, b = _problem(side, 1)
    A = A.astype(dtype)
    b = b.astype(dtype)
    x0 = np.zeros_like(b)
    eps = 1e-8 if dtype == np.float64 else 1e-5
    xo, used_o = oracle.cg_solver(A, b, x0, 5000, eps)
    D = H.DeviceEll.from_oracle(A)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray.from_host(x0)
    used, fd = H.cg_solve(D, db, dx, 5000, eps)
    x = dx.get()
    assert abs(used - used_o) <= (1 if dtype == np.float64 else 3)
    r0 = np.linalg.norm(b.astype(np.float64))
    r = np.linalg.norm(oracle.defect(A.astype(np.float64), b.astype(np.float64), x.astype(np.float64)))
    assert r < (eps * 1.5) * r0
    tol = 1e-8 if dtype == np.float64 else 2e-4
    assert np.abs(x - xo).max() <= tol * np.abs(xo).max()


def test_cg_q1_129_fp64_boundary_start():
    """Q1 129x129 problem: b = A x_true, x0 with Dirichlet values,
    FP64 CG to 1e-8 relative."""
    A, xt, b = _problem(129, 1)
    x0 = np.zeros_like(b)
    s = 129
    for i in range(s):
        x0[i], x0[-1 - i], x0[i * s] = b[i], b[-1 - i], b[i * s]
    D = H.DeviceEll.from_oracle(A)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray.from_host(x0)
    used, fd = H.cg_solve(D, db, dx, 10000, 1e-8)
    x = dx.get()
    xo, used_o = oracle.cg_solver(A, b, x0, 10000, 1e-8)
    assert abs(used - used_o) <= 1
    r = oracle.defect(A, b, x)
    assert np.linalg.norm(r) < 1e-8 * np.linalg.norm(oracle.defect(A, b, x0)) * 1.01
    assert np.abs(x - xt).max() < 1e-5


def test_cg_not_converged_reports():
    A, xt, b = _problem(65, 2)
    D = H.DeviceEll.from_oracle(A)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray(A.rows, np.float64)
    dx.zero()
    with pytest.raises(H.HipError) as e:
        H.cg_solve(D, db, dx, 3, 1e-12)
    assert e.value.code == 5


def _bf16_round(A):
    u = A.Ax.astype(np.float32).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    return oracle.Ell(A.rows, A.cols, A.stride, A.ncpr, A.threads, A.Aj, r, A.Arl)


def _pcg_case(side, vtype, threads, seed):
    dt = np.float64 if vtype == H.F64 else np.float32
    A = lognormal_q1_spd(side, seed, dt, threads)
    if vtype == H.BF16:
        A = _bf16_round(A)
    xt = uniform_pm1(seed + 1, A.rows).astype(dt)
    b = oracle.product(A, xt)
    x0 = uniform_pm1(seed + 2, A.rows).astype(dt)
    return A, b, x0, diag_inverse(A)


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("vtype", [H.F64, H.F32, H.BF16])
@pytest.mark.parametrize("side", [33, 129])
def test_pcg_jacobi_matches_oracle(side, vtype, threads):
    """ConjugateGradients<JAC> on a lognormal-coefficient SPD Q1 system from a
    random start: the iteration count of the oracle (reference control flow),
    the FP64 true residual below the tolerance and the iterate within
    rounding of the oracle's."""
    A, b, x0, dinv = _pcg_case(side, vtype, threads, 20 + side)
    eps = 1e-8 if vtype == H.F64 else 1e-5
    xo, used_o = oracle.pcg_jacobi(A, b, dinv, x0, 5000, eps)
    D = H.DeviceEll.from_oracle(A, vtype)
    db, dx, dd = H.DeviceArray.from_host(b), H.DeviceArray.from_host(x0), H.DeviceArray.from_host(dinv)
    print(f"pcg side={side} vtype={vtype} T={threads} oracle_iters={used_o}", flush=True)
    used, fd = H.pcg_jacobi(D, db, dx, dd, 5000, eps)
    x = dx.get()
    assert abs(used - used_o) <= (1 if vtype == H.F64 else 3)
    A64 = A.astype(np.float64)
    r0 = np.linalg.norm(oracle.defect(A64, b.astype(np.float64), x0.astype(np.float64)))
    r = np.linalg.norm(oracle.defect(A64, b.astype(np.float64), x.astype(np.float64)))
    assert r < (eps * (1.5 if vtype == H.F64 else 3.0)) * r0
    tol = 1e-7 if vtype == H.F64 else 1e-3
    assert np.abs(x - xo).max() <= tol * np.abs(xo).max()


def test_pcg_jacobi_area51_featflow():
    """math/dune_regression_TEST.cc:66-178 on the device: the area51 system
    from its start vector, FEATFLOW2's solution within 3e-5."""
    d = os.path.join(os.path.dirname(__file__), "golden", "honei_testdata", "l2")
    A = oracle.read_ell(os.path.join(d, "area51_full_0.ell"))
    b = oracle.read_exp(os.path.join(d, "area51_rhs_0"))
    x0 = oracle.read_exp(os.path.join(d, "area51_init_0"))
    sol = oracle.read_exp(os.path.join(d, "area51_sol_0"))
    dinv = diag_inverse(A)
    _xo, used_o = oracle.pcg_jacobi(A, b, dinv, x0, 10000, 1e-8)
    D = H.DeviceEll.from_oracle(A)
    db, dx, dd = H.DeviceArray.from_host(b), H.DeviceArray.from_host(x0), H.DeviceArray.from_host(dinv)
    used, _fd = H.pcg_jacobi(D, db, dx, dd, 10000, 1e-8)
    assert abs(used - used_o) <= 1
    assert np.abs(dx.get() - sol).max() < 3e-5


def test_pcg_jacobi_limits():
    """max_iters = 0: x untouched, 0 iterations; a capped run reports
    HONEI_HIP_ERROR_NOT_CONVERGED (5) with used == max_iters, or returns
    normally with check=False, as the reference does."""
    A, b, x0, dinv = _pcg_case(65, H.F64, 1, 7)
    D = H.DeviceEll.from_oracle(A)
    db, dd = H.DeviceArray.from_host(b), H.DeviceArray.from_host(dinv)
    dx = H.DeviceArray.from_host(x0)
    used, _ = H.pcg_jacobi(D, db, dx, dd, 0, 1e-8)
    assert used == 0 and np.array_equal(dx.get(), x0)
    with pytest.raises(H.HipError) as e:
        H.pcg_jacobi(D, db, dx, dd, 3, 1e-12)
    assert e.value.code == 5
    dx = H.DeviceArray.from_host(x0)
    used, _ = H.pcg_jacobi(D, db, dx, dd, 3, 1e-12, check=False)
    xo, used_o = oracle.pcg_jacobi(A, b, dinv, x0, 3, 1e-12)
    assert used == used_o == 3
    assert np.abs(dx.get() - xo).max() <= 1e-10 * np.abs(xo).max()


def test_pcg_jacobi_large_fp64_residual():
    """A 1025^2 lognormal SPD system (about 1M unknowns) to 1e-8: the FP64 true
    residual of the device result meets the tolerance (size-independent
    check; the oracle's run at this size is not needed for it)."""
    A, b, x0, dinv = _pcg_case(1025, H.F64, 1, 3)
    D = H.DeviceEll.from_oracle(A)
    db, dx, dd = H.DeviceArray.from_host(b), H.DeviceArray.from_host(x0), H.DeviceArray.from_host(dinv)
    used, fd = H.pcg_jacobi(D, db, dx, dd, 20000, 1e-8)
    r0 = np.linalg.norm(oracle.defect(A, b, x0))
    r = np.linalg.norm(oracle.defect(A, b, dx.get()))
    print(f"pcg 1025^2 iters={used} rel={r / r0:.3e}", flush=True)
    assert 0 < used < 20000 and r < 1.5e-8 * r0


@pytest.mark.parametrize("level", [2, 4, 6, 8])
def test_mg_fixed_fp64_matches_oracle(level):
    side = 2 ** level + 1
    A, xt, b = _problem(side, 3)
    Ho = oracle.q1_hierarchy(level, dtype=np.float64)
    cfg_o = oracle.mg_config(1, level, n_max_iter=50, tolerance=1e-10)
    xo, it_o, _ = oracle.mg_fixed(Ho, cfg_o, b)
    mg = H.Multigrid(H.F64, 1, level)
    mg.configure(H.mg_config(n_max_iter=50, tolerance=1e-10))
    mg.build_q1(0.7)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray(side * side, np.float64)
    dx.zero()
    it, fd = mg.solve_fixed(db, dx)
    x = dx.get()
    assert it == it_o
    assert np.abs(x - xo).max() <= 1e-9 * max(1.0, np.abs(xo).max())
    assert np.abs(x - xt).max() < 1e-8


@pytest.mark.parametrize("level,n_max_iter", [(3, 4), (4, 4), (5, 2), (6, 4), (8, 4)])
def test_mg_mixed_reference_semantics(level, n_max_iter):
    """Reference MIXED driver (inner start carried over between outer
    steps): same outer/inner counts as the CPU path, converged FP64 result."""
    side = 2 ** level + 1
    A, xt, b = _problem(side, 3)
    Ho = oracle.q1_hierarchy(level, dtype=np.float32)
    cfg_o = oracle.mg_config(1, level, n_max_iter=n_max_iter, tolerance=1e-8)
    xo, outer_o, inner_o, _ = oracle.mg_mixed(Ho, cfg_o, A, b)
    mg = H.Multigrid(H.F32, 1, level)
    mg.configure(H.mg_config(n_max_iter=n_max_iter, tolerance=1e-8))
    mg.build_q1(0.7)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray(side * side, np.float64)
    dx.zero()
    outer, inner, rd = mg.solve_mixed(db, dx)
    x = dx.get()
    assert (outer, inner) == (outer_o, inner_o)
    assert rd < 1e-8
    r = oracle.defect(A, b, x)
    assert np.linalg.norm(r) < 1e-8 * np.linalg.norm(oracle.defect(A, b, np.zeros_like(b)) ) * 10
    assert np.abs(x - xo).max() <= 1e-6 * np.abs(xo).max()
    assert np.abs(x - xt).max() < 1e-6


def test_mg_mixed_prolmat_hierarchy_from_host():
    """Hierarchy uploaded from host arrays with PROLMAT transfers (P and R =
    P^T as ELL products) -- the operator-list style data path."""
    level = 5
    side = 2 ** level + 1
    A, xt, b = _problem(side, 4)
    Ho = oracle.q1_hierarchy(level, dtype=np.float32, use_prolmat=True)
    cfg_o = oracle.mg_config(1, level, n_max_iter=4, tolerance=1e-8, prolmat=True)
    xo, outer_o, inner_o, _ = oracle.mg_mixed(Ho, cfg_o, A, b)
    mg = H.Multigrid(H.F32, 1, level)
    mg.configure(H.mg_config(n_max_iter=4, tolerance=1e-8, prolmat=True))
    for l in range(1, level + 1):
        mg.set_level(l, Ho.A[l], Ho.dinv[l])
        if l > 1:
            mg.set_transfer(l, 0, Ho.prol[l])
            mg.set_transfer(l, 1, Ho.res[l])
    mg.set_outer(A)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray(side * side, np.float64)
    dx.zero()
    outer, inner, rd = mg.solve_mixed(db, dx)
    assert (outer, inner) == (outer_o, inner_o)
    assert np.abs(dx.get() - xo).max() <= 1e-6 * np.abs(xo).max()


@pytest.mark.parametrize("level", [9, 11])
def test_mg_mixed_large_zero_start(level):
    """Q1 2049^2 shape (level 11, 4.2M unknowns): FP32 V(2,2) inner,
    FP64 outer to 1e-10; error vs x_true and vs an FP64 MG solve."""
    side = 2 ** level + 1
    n = side * side
    xt = uniform_pm1(3, n)
    A64 = oracle.q1_matrix(side)
    b = oracle.product(A64, xt)
    mg = H.Multigrid(H.F32, 1, level)
    mg.configure(H.mg_config(n_max_iter=1, tolerance=1e-10, zero_inner_start=True))
    mg.build_q1(0.7)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray(n, np.float64)
    dx.zero()
    outer, inner, rd = mg.solve_mixed(db, dx)
    x = dx.get()
    print(f"level {level}: outer {outer} inner {inner} rel defect {rd:.3e} err {np.abs(x - xt).max():.3e}")
    assert rd < 1e-10
    assert np.abs(x - xt).max() < 1e-8
    r = oracle.defect(A64, b, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-9


def _lognormal(seed, cells):
    return np.exp(np.random.default_rng(seed).standard_normal(cells))


@pytest.mark.parametrize("side", [65, 257])
def test_mg_mixed_lognormal_coefficient(side):
    """Variable-coefficient hierarchy (no reference counterpart): kappa per
    cell lognormal(0, 1), coarse kappa by geometric averaging of 2x2 cells,
    each level rediscretised.  Checked against the FP64 operator assembled
    on the device (itself pinned by brute-force element assembly): relative
    residual below 1e-8 and the error against x_true consistent with it."""
    n = side * side
    level = int(np.log2(side - 1))
    kappa = _lognormal(side, (side - 1) ** 2)
    Aj, Ax, Arl, _, stride = _device_q1(side, H.F64, coeff=kappa)
    A = oracle.Ell(n, n, stride, 9, 1, Aj.astype(np.uint64), Ax, Arl.astype(np.uint64))
    xt = uniform_pm1(side + 1, n)
    b = oracle.product(A, xt)
    mg = H.Multigrid(H.F32, 1, level)
    mg.configure(H.mg_config(n_max_iter=1, tolerance=1e-8, zero_inner_start=True, initial_zero=True,
                             max_outer=200))
    dk = H.DeviceArray.from_host(kappa, np.float64)
    mg.build_q1_coeff(dk, omega=0.8, averaging=1)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray(n, np.float64)
    dx.zero()
    outer, inner, rd = mg.solve_mixed(db, dx)
    x = dx.get()
    r = oracle.defect(A, b, x)
    print(f"side {side}: outer {outer} rel defect {rd:.3e} err {np.abs(x - xt).max():.3e}")
    assert outer < 40 and rd < 1e-8
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1.01e-8
    assert np.abs(x - xt).max() < 1e-4


def test_mg_coefficient_one_equals_constant_build():
    """kappa == 1 on every cell coarsens to 1 under every averaging and
    assembles the constant operator bit-for-bit, so the solve must be
    bit-identical to the constant-coefficient hierarchy."""
    side, level = 129, 7
    n = side * side
    b = oracle.product(oracle.q1_matrix(side), uniform_pm1(5, n))
    out = []
    for build in ("const", 0, 1, 2):
        mg = H.Multigrid(H.F32, 1, level)
        mg.configure(H.mg_config(n_max_iter=1, tolerance=1e-8, zero_inner_start=True, initial_zero=True))
        if build == "const":
            mg.build_q1(0.7)
        else:
            mg.build_q1_coeff(H.DeviceArray.from_host(np.ones((side - 1) ** 2), np.float64), 0.7, build)
        db, dx = H.DeviceArray.from_host(b), H.DeviceArray(n, np.float64)
        dx.zero()
        o, i, rd = mg.solve_mixed(db, dx)
        out.append((o, i, rd, dx.get()))
    for o, i, rd, x in out[1:]:
        assert (o, i, rd) == out[0][:3]
        assert np.array_equal(x, out[0][3])


@pytest.mark.parametrize("side,inner", [(65, "f64"), (129, "f32")])
def test_mg_pcg_lognormal(side, inner):
    """FP64 CG preconditioned by one V-cycle per iteration (no reference
    counterpart) on a lognormal-coefficient system with a rough right-hand
    side: converges, and the FP64 residual recomputed by the oracle with the
    device-assembled operator confirms the tolerance."""
    from tests.patterns import lognormal_cells
    n = side * side
    level = int(np.log2(side - 1))
    kappa = lognormal_cells(side + 3, (side - 1) ** 2)
    b = uniform_pm1(side + 4, n)
    Aj, Ax, Arl, _, stride = _device_q1(side, H.F64, coeff=kappa)
    A = oracle.Ell(n, n, stride, 9, 1, Aj.astype(np.uint64), Ax, Arl.astype(np.uint64))
    mg = H.Multigrid(H.F64 if inner == "f64" else H.F32, 1, level)
    mg.configure(H.mg_config(tolerance=1e-8, max_outer=100))
    mg.build_q1_coeff(H.DeviceArray.from_host(kappa, np.float64), 0.8, 1)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray(n, np.float64)
    dx.zero()
    it, rd = mg.solve_pcg(db, dx)
    x = dx.get()
    bd = np.zeros(n, bool)
    bd[:side] = bd[-side:] = True
    bd[::side] = True
    r0 = np.linalg.norm(oracle.defect(A, b, np.where(bd, b, 0.0)))
    r = np.linalg.norm(oracle.defect(A, b, x))
    print(f"side {side} {inner}: {it} iterations, rel {rd:.2e}, host rel {r / r0:.2e}")
    assert rd < 1e-8 and r / r0 < 1.001e-8 and it < 40
