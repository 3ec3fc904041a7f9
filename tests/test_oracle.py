"""Pin the CPU oracle before trusting it.

Checks, in order of strength:
  * the reference's own golden files (bit-exact CSR->ELL, FEAT solution);
  * the reference itself (oracle/_ref, HONEI's tags::CPU code built from its
    sources) -- bit-exact, skipped where it was not built;
  * mathematics: dense brute force, closed forms of the Q1 stencil and of
    bilinear transfers, invariants (geometric transfer == PROLMAT product).
"""
import os

import numpy as np
import pytest

import oracle
from oracle import ref
from tests.patterns import honei_pattern_csr, random_band_csr, uniform_pm1
from tests.spd import diag_inverse, lognormal_q1_spd

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference sources absent)")


# ---------------------------------------------------------------------------
# CSR -> ELL (index arrays must be bit-exact)
# ---------------------------------------------------------------------------
def test_mtx_to_ell_matches_reference_file(golden):
    """testdata/5pt_10x10.ell was written by clients/io_tools/mtxtoell.cc from
    5pt_10x10.mtx via SparseMatrix -> SparseMatrixELL(T=1)."""
    rows, cols, rp, col, val = oracle.read_mtx(os.path.join(golden, "5pt_10x10.mtx"))
    A = oracle.ell_from_sm(rows, cols, rp, col, val, threads=1)
    F = oracle.read_ell(os.path.join(golden, "5pt_10x10.ell"))
    assert (A.stride, A.ncpr) == (F.stride, F.ncpr) == (128, 5)
    assert np.array_equal(A.Aj, F.Aj)
    assert np.array_equal(A.Ax.view(np.uint64), F.Ax.view(np.uint64))
    assert np.array_equal(A.Arl, F.Arl)
    assert A.used_elements() == 460


@pytest.mark.parametrize("name", [f"poisson/A_{i}.ell" for i in range(1, 7)]
                         + [f"poisson/prol_{i}.ell" for i in range(2, 7)]
                         + [f"poisson_advanced4/sort_0/A_{i}.ell" for i in range(1, 5)])
def test_reference_ell_files_round_trip(golden, name):
    """Every fixture was written through SparseMatrix -> ELL(T=1)
    (clients/io_tools/feasttoell.cc:128-132), so re-converting its contents
    must reproduce the file bit-for-bit."""
    F = oracle.read_ell(os.path.join(golden, name))
    rp, col, val = F.to_csr()
    A = oracle.ell_from_sm(F.rows, F.cols, rp, col, val, threads=1)
    assert (A.stride, A.ncpr) == (F.stride, F.ncpr)
    assert np.array_equal(A.Aj, F.Aj)
    assert np.array_equal(A.Ax.view(np.uint64), F.Ax.view(np.uint64))
    assert np.array_equal(A.Arl, F.Arl)


@pytest.mark.parametrize("threads", [1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ell_threads_layout_closed_form(threads, dtype):
    """la/sparse_matrix_ell_TEST.cc:45-70 pattern: slot of the t-th non-zero of
    row r is (t % T) + r*T + (t / T)*stride, stride = 32*ceil(rows*T/32),
    Arl = ceil(nnz_r / T); SparseMatrix(ELL) gives the input back."""
    rows, cols = 111, 114
    rp, col, val = honei_pattern_csr(rows, cols, plus_one=True, corner=4711, dtype=dtype)
    A = oracle.ell_from_sm(rows, cols, rp, col, val, threads)
    lens = np.diff(rp.astype(np.int64))
    assert A.stride == 32 * ((rows * threads + 31) // 32)
    assert A.ncpr == int(np.ceil(max(1, lens.max()) / threads))
    assert np.array_equal(A.Arl, np.ceil(lens / threads).astype(np.uint64))
    for r in range(rows):
        for t in range(int(lens[r])):
            s = (t % threads) + r * threads + (t // threads) * A.stride
            assert A.Aj[s] == col[rp[r] + t] and A.Ax[s] == val[rp[r] + t]
    assert A.used_elements() == len(val)
    rp2, col2, val2 = A.to_csr()
    assert np.array_equal(rp2, rp) and np.array_equal(col2, col) and np.array_equal(val2, val)
    # CSR constructor path gives the same matrix (la/sparse_matrix_ell-impl.hh:124-183)
    B = oracle.ell_from_csr(rows, cols, rp, col, val, threads)
    assert np.array_equal(A.Aj, B.Aj) and np.array_equal(A.Ax, B.Ax) and np.array_equal(A.Arl, B.Arl)


@needs_ref
@pytest.mark.parametrize("threads", [1, 2, 4, 8, 16, 32])
def test_ell_layout_matches_reference_build(threads):
    rp, col, val = random_band_csr(2000, seed=7, max_len=27, band=300)
    A = oracle.ell_from_sm(2000, 2000, rp, col, val, threads)
    R = ref.ell_from_sm(2000, 2000, rp, col, val, threads).to_ell()
    assert (A.stride, A.ncpr, A.threads) == (R.stride, R.ncpr, R.threads)
    assert np.array_equal(A.Aj, R.Aj) and np.array_equal(A.Ax, R.Ax) and np.array_equal(A.Arl, R.Arl)
    Rc = ref.ell_from_csr(2000, 2000, rp, col, val, threads).to_ell()
    assert np.array_equal(A.Aj, Rc.Aj) and np.array_equal(A.Ax, Rc.Ax) and np.array_equal(A.Arl, Rc.Arl)


def test_ell_row_length_from_arrays():
    """Explicit-array constructor derives Arl from Ax (la/sparse_matrix_ell-impl.hh:186-208)."""
    rp, col, val = random_band_csr(300, seed=3, max_len=9, band=20)
    for T in (1, 2, 4):
        A = oracle.ell_from_sm(300, 300, rp, col, val, T)
        B = oracle.ell_from_arrays(300, 300, A.stride, A.ncpr, A.Aj, A.Ax, T)
        assert np.array_equal(A.Arl, B.Arl)


# ---------------------------------------------------------------------------
# Products
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("threads", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("size", [11, 22, 44, 88, 176, 352])
def test_product_vs_dense(size, threads):
    """la/product_TEST.cc:786-803: size x (size+3), x = 4 except x0=1, x1=2."""
    rp, col, val = honei_pattern_csr(size, size + 3)
    A = oracle.ell_from_sm(size, size + 3, rp, col, val, threads)
    x = np.full(size + 3, 4.0)
    x[0], x[1] = 1, 2
    y = oracle.product(A, x)
    D = np.zeros((size, size + 3))
    for r in range(size):
        D[r, col[rp[r]:rp[r + 1]].astype(int)] = val[rp[r]:rp[r + 1]]
    np.testing.assert_allclose(y, D @ x, rtol=1e-13, atol=1e-9)


def test_product_mc_equals_generic():
    """mc::Product splits rows only; per-row arithmetic is identical."""
    rp, col, val = random_band_csr(5000, seed=11)
    A = oracle.ell_from_sm(5000, 5000, rp, col, val, 2)
    x = uniform_pm1(5, 5000)
    assert np.array_equal(oracle.product(A, x), oracle.product_mc(A, x, 7))


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("threads", [1, 2, 4, 8, 16])
def test_product_bitexact_vs_reference(dtype, threads):
    rp, col, val = random_band_csr(3000, seed=13, max_len=27, band=500, dtype=dtype)
    x = uniform_pm1(2, 3000).astype(dtype)
    A = oracle.ell_from_sm(3000, 3000, rp, col, val, threads)
    R = ref.ell_from_sm(3000, 3000, rp, col, val, threads)
    assert np.array_equal(oracle.product(A, x).view(np.uint8), R.product(x).view(np.uint8))
    b = uniform_pm1(9, 3000).astype(dtype)
    assert np.array_equal(oracle.defect(A, b, x).view(np.uint8), R.defect(b, x).view(np.uint8))


def test_defect_5pt(golden):
    """math/defect_TEST.cc:45-72: 5pt_10x10, x_i = i/1.234, rhs = 1.234."""
    A = oracle.read_ell(os.path.join(golden, "5pt_10x10.ell"))
    x = np.arange(100) / 1.234
    rhs = np.full(100, 1.234)
    r = oracle.defect(A, rhs, x)
    np.testing.assert_allclose(r, rhs - A.to_dense() @ x, rtol=1e-13, atol=1e-11)


def test_blas1_closed_forms():
    x = uniform_pm1(1, 1001)
    y = uniform_pm1(2, 1001)
    assert oracle.dot(x, y) == pytest.approx(float(x @ y), rel=1e-13)
    assert oracle.norm2(x) == pytest.approx(float(x @ x), rel=1e-13)
    assert oracle.norm2(x, root=True) == pytest.approx(float(np.sqrt(x @ x)), rel=1e-13)
    np.testing.assert_array_equal(oracle.scaled_sum2(x, y, 0.5), x + 0.5 * y)
    np.testing.assert_array_equal(oracle.scaled_sum3(x, y, -2.0), x + -2.0 * y)
    np.testing.assert_array_equal(oracle.scaled_sum_v(x, y, x), x + y * x)


# ---------------------------------------------------------------------------
# Q1 operator and transfers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("side", [3, 5, 9, 17, 129])
def test_q1_closed_form(side):
    """math/fill_matrix.hh:36-78: identity Dirichlet rows on the boundary,
    8/3 on the diagonal and -1/3 on the 8 neighbours inside."""
    A = oracle.q1_matrix(side)
    n = side * side
    inner = (side - 2) ** 2
    assert A.used_elements() == 9 * inner + (n - inner)
    assert A.ncpr == (9 if side > 2 else 1) and A.stride == 32 * ((n + 31) // 32)
    if side <= 17:
        D = A.to_dense()
        for r in range(n):
            i, j = divmod(r, side)
            if i in (0, side - 1) or j in (0, side - 1):
                assert D[r, r] == 1 and np.count_nonzero(D[r]) == 1
            else:
                assert D[r, r] == pytest.approx(8 / 3)
                for di in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        if di or dj:
                            assert D[r, (i + di) * side + j + dj] == pytest.approx(-1 / 3)
                assert abs(D[r].sum()) < 1e-15


@needs_ref
@pytest.mark.parametrize("bc", [0, 1])
def test_q1_matches_reference_build(bc):
    for dt in (np.float32, np.float64):
        A = oracle.q1_matrix(33, dt, bc=bc)
        R = ref.q1_ell(33, dt, bc=bc).to_ell()
        assert np.array_equal(A.Aj, R.Aj) and np.array_equal(A.Ax, R.Ax) and np.array_equal(A.Arl, R.Arl)


def test_prolongation_matrix_known_answer():
    """math/prolongation_matrix_TEST.cc:40-80: the 25x9 matrix (fine side 5)."""
    P = oracle.prolmat(5).to_dense()
    assert P.shape == (25, 9)
    for i in range(5):
        for j in range(5):
            row = P[i * 5 + j]
            expect = np.zeros((3, 3))
            for ci in range(3):
                for cj in range(3):
                    wi = max(0.0, 1 - abs(i - 2 * ci) / 2)
                    wj = max(0.0, 1 - abs(j - 2 * cj) / 2)
                    expect[ci, cj] = wi * wj
            np.testing.assert_array_equal(row, expect.ravel())


@needs_ref
def test_prolongation_matrix_matches_reference_build():
    for side in (5, 9, 33):
        A = oracle.prolmat(side)
        R = ref.prolmat(side).to_ell()
        assert np.array_equal(A.Aj, R.Aj) and np.array_equal(A.Ax, R.Ax) and np.array_equal(A.Arl, R.Arl)


@pytest.mark.parametrize("level", [1, 2, 3, 5, 7])
def test_geometric_transfers_equal_prolmat(level):
    """math/prolongation_TEST.cc:93-138: geometric == PROLMAT + Dirichlet; the
    restriction is the transpose of the same matrix."""
    side_c = 2 ** level + 1
    side_f = 2 ** (level + 1) + 1
    P = oracle.prolmat(side_f)
    c = uniform_pm1(level, side_c * side_c)
    f = uniform_pm1(level + 100, side_f * side_f)
    np.testing.assert_allclose(oracle.prolongation(c), oracle.apply_dirichlet(oracle.product(P, c)),
                               rtol=0, atol=1e-15)
    R = oracle.transpose(P)
    np.testing.assert_allclose(oracle.restriction(f), oracle.apply_dirichlet(oracle.product(R, f)),
                               rtol=0, atol=1e-14)


@pytest.mark.parametrize("level", list(range(1, 8)))
def test_restriction_of_constant(level):
    """math/restriction_TEST.cc:45-77: a constant fine field restricts to 4x
    inside (1 + 4*1/2 + 4*1/4) and to 0 on Dirichlet nodes."""
    side_f = 2 ** (level + 1) + 1
    c = oracle.restriction(np.ones(side_f * side_f, np.float32)).reshape(2 ** level + 1, -1)
    assert (c[1:-1, 1:-1] == 4).all()
    assert (c[0] == 0).all() and (c[-1] == 0).all() and (c[:, 0] == 0).all() and (c[:, -1] == 0).all()


@needs_ref
@pytest.mark.parametrize("level", [1, 2, 4, 6])
def test_transfers_bitexact_vs_reference(level):
    side_c = 2 ** level + 1
    side_f = 2 ** (level + 1) + 1
    c = uniform_pm1(3, side_c * side_c).astype(np.float32)
    f = uniform_pm1(4, side_f * side_f).astype(np.float32)
    for mask in [(2,) * 8, (2, 2, 2, 2, 2, 1, 2, 2), (1,) * 8]:
        assert np.array_equal(oracle.prolongation(c, mask), ref.prolongation(c, mask))
        assert np.array_equal(oracle.restriction(f, mask), ref.restriction(f, mask))


# ---------------------------------------------------------------------------
# Solvers
# ---------------------------------------------------------------------------
def _q1_problem(side, seed):
    A = oracle.q1_matrix(side)
    xt = uniform_pm1(seed, side * side)
    return A, xt, oracle.product(A, xt)


def test_cg_solver_closed_form():
    """CGSolver to 1e-8 relative on b = A x_true (the Q1 CG workload, smaller)."""
    A, xt, b = _q1_problem(33, 1)
    x0 = np.zeros_like(b)
    s = 33
    for i in range(s):
        x0[i], x0[-1 - i], x0[i * s] = b[i], b[-1 - i], b[i * s]
    x, used, hist = oracle.cg_solver(A, b, x0, 5000, 1e-8, history=True)
    r = oracle.defect(A, b, x)
    assert np.linalg.norm(r) < 1e-8 * np.linalg.norm(oracle.defect(A, b, x0))
    assert 0 < used < 5000 and hist[-1] < hist[0]
    assert np.abs(x - xt).max() < 1e-5


@needs_ref
def test_cg_solver_bitexact_vs_reference():
    A, _xt, b = _q1_problem(65, 1)
    x0 = np.zeros_like(b)
    x, used = oracle.cg_solver(A, b, x0, 3000, 1e-8)
    R = ref.q1_ell(65)
    xr, used_r = ref.cg_solver(R, b, x0, 3000, 1e-8)
    assert used == used_r
    assert np.array_equal(x, xr)


def test_pcg_jacobi_area51_featflow(golden):
    """math/dune_regression_TEST.cc:66-178: PCG-Jacobi to 1e-8 from the stored
    start vector reproduces FEATFLOW2's solution within that test's eps
    (3 * (m*8 + b) = 3e-5 for double)."""
    d = os.path.join(golden, "l2")
    A = oracle.read_ell(os.path.join(d, "area51_full_0.ell"))
    b = oracle.read_exp(os.path.join(d, "area51_rhs_0"))
    x0 = oracle.read_exp(os.path.join(d, "area51_init_0"))
    sol = oracle.read_exp(os.path.join(d, "area51_sol_0"))
    x, used = oracle.pcg_jacobi(A, b, diag_inverse(A), x0, 10000, 1e-8)
    assert 0 < used < 10000
    assert np.abs(x - sol).max() < 3e-5


def test_pcg_jacobi_diagonal_system_one_iteration():
    """With A diagonal and dinv = 1/diag(A), z_0 is the exact error: one
    iteration solves the system (step = 1, r_1 = 0).  Without the
    preconditioner CG needs one iteration per distinct diagonal value, so a
    dropped or misplaced dinv fails here."""
    n = 40
    diag = np.repeat(np.array([1.0, 2.0, 5.0, 11.0]), n // 4)
    rp = np.arange(n + 1, dtype=np.uint64)
    A = oracle.ell_from_sm(n, n, rp, np.arange(n, dtype=np.uint64), diag)
    b = uniform_pm1(3, n)
    x, used = oracle.pcg_jacobi(A, b, 1.0 / diag, None, 100, 1e-12)
    assert used == 1
    np.testing.assert_allclose(x, b / diag, rtol=4e-16, atol=0)
    _x, used_plain = oracle.pcg_jacobi(A, b, np.ones(n), None, 100, 1e-12)
    assert used_plain == 4


def test_pcg_jacobi_unit_preconditioner_is_plain_cg():
    """dinv = 1 makes z = r exactly, and every PCG-Jacobi quantity then equals
    CGSolver's (r.z = ||r||^2, the same step, the same direction update with
    operands commuted): iterates and defect histories agree bit-for-bit with
    the independently restated CGSolver."""
    A = lognormal_q1_spd(33, 4, np.float64)
    b = uniform_pm1(5, A.rows)
    xp, used_p, hp = oracle.pcg_jacobi(A, b, np.ones(A.rows), None, 3000, 1e-8, history=True)
    xc, used_c, hc = oracle.cg_solver(A, b, None, 3000, 1e-8, history=True)
    assert used_p == used_c > 10
    assert np.array_equal(xp, xc) and np.array_equal(hp, hc)


def test_pcg_jacobi_limits():
    """max_iters = 0 leaves x and used_iters alone; a capped run reports
    max_iters and is a prefix of the full run; the preconditioned solve of a
    lognormal-coefficient system needs fewer iterations than plain CG."""
    A = lognormal_q1_spd(33, 6, np.float64, sigma=1.5)
    b = uniform_pm1(7, A.rows)
    dinv = diag_inverse(A)
    x0 = uniform_pm1(8, A.rows)
    x, used = oracle.pcg_jacobi(A, b, dinv, x0, 0, 1e-8)
    assert used == 0 and np.array_equal(x, x0)
    xf, used_f, hf = oracle.pcg_jacobi(A, b, dinv, x0, 5000, 1e-8, history=True)
    x3, used_3, h3 = oracle.pcg_jacobi(A, b, dinv, x0, 3, 1e-8, history=True)
    assert used_3 == 3 and np.array_equal(h3, hf[:3])
    assert np.linalg.norm(oracle.defect(A, b, xf)) < 1e-8 * np.linalg.norm(oracle.defect(A, b, x0)) * 1.01
    _xc, used_c = oracle.cg_solver(A, b, x0, 5000, 1e-8)
    assert used_f < used_c


@needs_ref
@pytest.mark.parametrize("dtype,threads", [(np.float32, 1), (np.float64, 1), (np.float64, 4)])
def test_pcg_jacobi_bitexact_vs_reference(dtype, threads):
    A = lognormal_q1_spd(33, 9, dtype, threads)
    rp, col, val = A.to_csr()
    R = ref.ell_from_sm(A.rows, A.cols, rp, col, val, threads)
    b = uniform_pm1(10, A.rows).astype(dtype)
    x0 = uniform_pm1(11, A.rows).astype(dtype)
    dinv = diag_inverse(A)
    x, used = oracle.pcg_jacobi(A, b, dinv, x0, 2000, 1e-6 if dtype == np.float32 else 1e-8)
    xr, used_r = ref.pcg_jacobi(R, b, dinv, x0, 2000, 1e-6 if dtype == np.float32 else 1e-8)
    assert used == used_r > 0
    assert np.array_equal(x, xr)


@needs_ref
def test_pcg_jacobi_area51_bitexact_vs_reference(golden):
    d = os.path.join(golden, "l2")
    A = oracle.read_ell(os.path.join(d, "area51_full_0.ell"))
    rp, col, val = A.to_csr()
    R = ref.ell_from_sm(A.rows, A.cols, rp, col, val, 1)
    b = oracle.read_exp(os.path.join(d, "area51_rhs_0"))
    x0 = oracle.read_exp(os.path.join(d, "area51_init_0"))
    dinv = diag_inverse(A)
    for its in (0, 2, 10000):
        x, used = oracle.pcg_jacobi(A, b, dinv, x0, its, 1e-8)
        xr, used_r = ref.pcg_jacobi(R, b, dinv, x0, its, 1e-8)
        assert used == used_r and np.array_equal(x, xr)


@needs_ref
def test_cg_konv_bitexact_vs_reference():
    A = oracle.q1_matrix(9, np.float32)
    b = uniform_pm1(2, 81).astype(np.float32)
    x, it = oracle.cg_konv(A, b)
    xr = ref.cg_konv(ref.q1_ell(9, np.float32), b, float(np.finfo(np.float32).eps))
    assert it > 0 and np.array_equal(x, xr)


@pytest.mark.parametrize("level", [2, 4, 6])
def test_mg_fixed_fp64_converges(level):
    side = 2 ** level + 1
    A, xt, b = _q1_problem(side, 3)
    H = oracle.q1_hierarchy(level, dtype=np.float64)
    cfg = oracle.mg_config(1, level, n_max_iter=50, tolerance=1e-10)
    x, it, fd = oracle.mg_fixed(H, cfg, b)
    assert np.linalg.norm(oracle.defect(A, b, x)) < 1e-9 * np.linalg.norm(b)
    assert np.abs(x - xt).max() < 1e-8


@needs_ref
@pytest.mark.parametrize("level,n_max_iter", [(3, 4), (4, 4), (5, 2), (6, 4)])
def test_mg_mixed_bitexact_vs_reference(level, n_max_iter):
    """Whole Multigrid<..., JAC, V, MIXED> solve: same iteration counts and a
    bit-identical FP64 solution as HONEI's own code."""
    side = 2 ** level + 1
    A, xt, b = _q1_problem(side, 3)
    H = oracle.q1_hierarchy(level, dtype=np.float32)
    cfg = oracle.mg_config(1, level, n_max_iter=n_max_iter, tolerance=1e-8)
    x, outer, inner, hist = oracle.mg_mixed(H, cfg, A, b)
    xr, outer_r, inner_r = ref.mg_mixed(level, b, np.zeros_like(b), n_max_iter=n_max_iter)
    assert (outer, inner) == (outer_r, inner_r)
    assert np.array_equal(x, xr)
    assert hist[-1] < 1e-8


def test_mg_mixed_reaches_fp64_accuracy():
    level = 6
    A, xt, b = _q1_problem(2 ** level + 1, 4)
    H = oracle.q1_hierarchy(level, dtype=np.float32)
    cfg = oracle.mg_config(1, level, n_max_iter=4, tolerance=1e-12)
    x, outer, inner, hist = oracle.mg_mixed(H, cfg, A, b)
    assert hist[-1] < 1e-12 and np.abs(x - xt).max() < 1e-10
    cfg0 = oracle.mg_config(1, level, n_max_iter=4, tolerance=1e-12, zero_inner_start=True)
    x0, outer0, _, hist0 = oracle.mg_mixed(H, cfg0, A, b)
    assert hist0[-1] < 1e-12 and np.abs(x0 - xt).max() < 1e-10


def test_mg_oplist_w_cycle_vs_feat_solution(golden):
    """math/mg_TEST.cc:389-464: W-cycle MG (CG coarse, Jacobi/RI smoother with
    damping 0.7, 4+4 sweeps, PROLMAT) on testdata/poisson L1..L6 vs sol_6."""
    base = os.path.join(golden, "poisson")
    H = oracle.MGHierarchy()
    for l in range(1, 7):
        A = oracle.read_ell(os.path.join(base, f"A_{l}.ell"))
        H.A[l] = A
        D = _diag(A)
        H.dinv[l] = (0.7 / D).astype(np.float64)
        if l >= 2:
            P = oracle.read_ell(os.path.join(base, f"prol_{l}.ell"))
            H.prol[l] = P
            H.res[l] = oracle.transpose(P)
    b = oracle.read_exp(os.path.join(base, "rhs_6"))
    x0 = oracle.read_exp(os.path.join(base, "init_6"))
    sol = oracle.read_exp(os.path.join(base, "sol_6"))
    x, used, norms = oracle.mg_oplist(H, 1, 6, b, x0, cycle="W", max_iters=100, max_iters_coarse=100,
                                      n_pre=4, n_post=4, eps=1e-8)
    # tolerance derived as in math/mg_TEST.cc:437-446
    base_eps, add_eps = 1e-1, 1e-3
    m = (add_eps - base_eps) / 4
    eps = (m * 8 + (base_eps - 4 * m)) * 8
    assert norms[1] < 1e-8 * norms[0]
    assert np.abs(x - sol).max() < eps
    # and much tighter in fact: both are converged solutions of the same system
    assert np.abs(x - sol).max() < 1e-5


def _diag(A):
    d = np.zeros(A.rows)
    for r in range(A.rows):
        for k in range(int(A.Arl[r])):
            for t in range(A.threads):
                s = r * A.threads + t + k * A.stride
                if int(A.Aj[s]) == r and A.Ax[s] != 0:
                    d[r] = A.Ax[s]
    return d


# ---- variable-coefficient Q1 (no reference counterpart) ------------------------------

def _element_assembly(side, kappa):
    """Brute force: bilinear Q1 element stiffness per unit cell, summed."""
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


@pytest.mark.parametrize("side", [3, 5, 9])
def test_q1_coeff_matrix_element_assembly(side):
    from tests.patterns import lognormal_cells
    kappa = lognormal_cells(side, (side - 1) ** 2)
    A = oracle.q1_coeff_matrix(side, kappa)
    np.testing.assert_allclose(A.to_dense(), _element_assembly(side, kappa), rtol=1e-14, atol=1e-15)
    # kappa = 1: the reference's constant operator, bit for bit
    for dt in (np.float32, np.float64):
        A1 = oracle.q1_coeff_matrix(side, np.ones((side - 1) ** 2), dt)
        R = oracle.q1_matrix(side, dt)
        assert np.array_equal(A1.Ax, R.Ax) and np.array_equal(A1.Aj, R.Aj) and np.array_equal(A1.Arl, R.Arl)


def test_coarsen_coeff_closed_forms():
    rng = np.random.default_rng(0)
    k = np.exp(rng.standard_normal(64))
    f = k.reshape(8, 8)
    blocks = np.stack([f[0::2, 0::2], f[0::2, 1::2], f[1::2, 0::2], f[1::2, 1::2]])
    np.testing.assert_allclose(oracle.coarsen_coeff(k, 8, 0).reshape(4, 4), blocks.mean(0), rtol=1e-15)
    np.testing.assert_allclose(oracle.coarsen_coeff(k, 8, 1).reshape(4, 4), np.exp(np.log(blocks).mean(0)),
                               rtol=1e-14)
    np.testing.assert_allclose(oracle.coarsen_coeff(k, 8, 2).reshape(4, 4), 4.0 / (1.0 / blocks).sum(0),
                               rtol=1e-15)
    c = np.full(64, 2.5)
    for avg in (0, 1, 2):
        np.testing.assert_allclose(oracle.coarsen_coeff(c, 8, avg), 2.5, rtol=1e-15)


def test_mg_mixed_lognormal_hierarchy_converges():
    """MIXED defect correction on a rediscretised lognormal hierarchy with a
    rough right-hand side reaches 1e-8 (FP64 residual recomputed here)."""
    from tests.patterns import lognormal_cells
    L = 5
    side = 2 ** L + 1
    kappa = lognormal_cells(31, (side - 1) ** 2)
    b = uniform_pm1(32, side * side)
    Hh = oracle.q1_coeff_hierarchy(L, kappa, averaging=1, omega=0.8)
    A64 = oracle.q1_coeff_matrix(side, kappa)
    cfg = oracle.mg_config(1, L, n_max_iter=1, tolerance=1e-8, initial_zero=True, zero_inner_start=True,
                           max_outer=60)
    x, outer, inner, hist = oracle.mg_mixed(Hh, cfg, A64, b)
    assert outer < 40 and hist[-1] < 1e-8
    bd = np.zeros(side * side, bool)
    bd[:side] = bd[-side:] = True
    bd[::side] = True
    r0 = np.linalg.norm(oracle.defect(A64, b, np.where(bd, b, 0.0)))
    assert np.linalg.norm(oracle.defect(A64, b, x)) / r0 < 1.001e-8


# ---- banded Q1 product / defect (la/product.hh:371-400, math/defect.hh:194-199) ------

@pytest.mark.parametrize("side", [3, 5, 33])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_banded_q1_vs_reference_and_ell(side, dtype):
    from tests.patterns import lognormal_cells
    x = uniform_pm1(side, side * side).astype(dtype)
    rhs = uniform_pm1(side + 7, side * side).astype(dtype)
    for bands in (oracle.q1_bands(side, dtype), oracle.q1_coeff_bands(side, lognormal_cells(side, (side - 1) ** 2),
                                                                       dtype)):
        y = oracle.banded_q1(bands, x)
        r = oracle.banded_q1(bands, x, rhs)
        if ref.available():
            assert np.array_equal(y.view(np.uint8), ref.banded_q1(bands, x).view(np.uint8))
            assert np.array_equal(r.view(np.uint8), ref.banded_q1(bands, x, rhs).view(np.uint8))
        # same operator as the ELL conversion, different summation order
        rp, col, val = oracle.banded_to_csr(bands)
        E = oracle.ell_from_sm(side * side, side * side, rp, col, val)
        ye = oracle.product(E, x)
        tol = 4 * np.finfo(dtype).eps * (np.abs(bands).sum(0) * np.abs(x).max() + 1e-30)
        assert np.all(np.abs(y - ye) <= tol)
        assert np.array_equal(r, (rhs - y).astype(dtype))


# ---------------------------------------------------------------------------
# FP64 direct solve (the error reference of the solver results)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("threads", [1, 4])
def test_csr_arrays_equal_sparse_matrix_rows(golden, threads):
    """The vectorised CSR extraction gives the rows SparseMatrix(ELL) gives
    (la/sparse_matrix.hh:157-185): on the reference's 5pt_10x10 file and on
    a threads-per-row layout with explicit zeros stored."""
    F = oracle.read_ell(os.path.join(golden, "5pt_10x10.ell"))
    kappa = np.exp(uniform_pm1(3, 64))
    B = oracle.q1_coeff_matrix(9, kappa, threads=threads)
    B.Ax[B.Aj == 40] = 0.0   # stored zeros are not entries
    for A in (F, B):
        rp, c, v = A.to_csr()
        rp2, c2, v2 = oracle.csr_arrays(A)
        assert np.array_equal(rp.astype(np.int64), rp2)
        assert np.array_equal(c.astype(np.int64), c2)
        assert np.array_equal(v, v2)


@pytest.mark.parametrize("side", [17, 65])
def test_direct_solve_recovers_known_solution(side):
    """b = A x_true on the constant and the lognormal-coefficient Q1
    operators: the direct solve gives x_true back to rounding, and its own
    residual (the oracle's Defect) is at machine precision."""
    from tests.patterns import lognormal_cells
    for A in (oracle.q1_matrix(side), oracle.q1_coeff_matrix(side, lognormal_cells(side, (side - 1) ** 2))):
        xt = uniform_pm1(side + 1, side * side)
        b = oracle.product(A, xt)
        x, rel, steps = oracle.direct_solve(A, b)
        assert rel <= 1e-14 and steps <= 4
        assert np.abs(x - xt).max() <= 1e-11


def test_direct_solve_agrees_with_cg_solver():
    """CGSolver (math/cg.hh:63-133) run to 1e-13 and the direct solve agree
    to the accuracy the CG tolerance implies; a dropped or transposed term in
    either would separate them far beyond it."""
    A = oracle.q1_matrix(33)
    b = uniform_pm1(44, A.rows)
    xc, used = oracle.cg_solver(A, b, np.zeros_like(b), 10000, 1e-13)
    xd, rel, _ = oracle.direct_solve(A, b)
    assert rel <= 1e-14
    assert np.abs(xc - xd).max() <= 1e-8 * np.abs(xd).max()   # measured 1.0e-9
    # the transpose solves a different system (the Dirichlet rows make A non-symmetric)
    xt_, _, _ = oracle.direct_solve(oracle.transpose(A), b)
    assert np.abs(xt_ - xd).max() > 1e-3 * np.abs(xd).max()
