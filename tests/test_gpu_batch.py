"""GPU tests of the batched multigrid solver (honei_hip_batch_*).

The batch has no reference counterpart; its per-system numerics are those of
the single-system MIXED driver (math/multigrid.hh:2395-2536 as restated by
MGHierarchy::solve_mixed), so the bars are:
  * threads-per-row 1 (autotune off): every slot's solution, outer iteration
    count and relative defect are bit-identical to the single-system solve
    of the same system;
  * a system's result does not depend on the other systems in the batch or
    on slot reuse (bit-identical alone and in the mix, with autotuning on);
  * each solution's FP64 residual, computed on the host by the oracle with
    the device-assembled FP64 operator, is below the tolerance.
"""
import numpy as np
import pytest

import honei_amd as H
import oracle
from tests.patterns import lognormal_cells, poisson_system_seeds, uniform01, uniform_pm1

pytestmark = pytest.mark.gpu

OMEGA, AVG, TOL = 0.8, 1, 1e-8


@pytest.fixture(scope="module", autouse=True)
def _init():
    H.init(0)
    yield


def _system(level, i, base=11):
    side = 2 ** level + 1
    cs, bs = poisson_system_seeds(base, i)
    return lognormal_cells(cs, (side - 1) ** 2), uniform_pm1(bs, side * side)


def _fp64_operator(side, kappa):
    n = side * side
    stride = 32 * ((n + 31) // 32)
    Aj = H.DeviceArray(9 * stride, np.int32)
    Ax = H.DeviceArray(9 * stride, np.float64)
    Arl = H.DeviceArray(n, np.uint32)
    dk = H.DeviceArray.from_host(kappa, np.float64)
    H.call("honei_hip_q1_assemble", H.F64, side, Aj.ptr, Ax.ptr, Arl.ptr, stride, dk.ptr, None, 0.0, 1, None)
    return oracle.Ell(n, n, stride, 9, 1, Aj.get().astype(np.uint64), Ax.get(), Arl.get().astype(np.uint64))


def _single(level, kappa, b, inner=H.F32):
    n = (2 ** level + 1) ** 2
    mg = H.Multigrid(inner, 1, level)
    mg.configure(H.mg_config(n_max_iter=1, tolerance=TOL, zero_inner_start=True, initial_zero=True,
                             max_outer=200))
    mg.build_q1_coeff(H.DeviceArray.from_host(kappa, np.float64), OMEGA, AVG)
    db, dx = H.DeviceArray.from_host(b), H.DeviceArray(n, np.float64)
    dx.zero()
    outer, _, rd = mg.solve_mixed(db, dx, check=False)
    return outer, rd, dx.get()


def _run_batch(levels, counts, systems, **cfg):
    """systems: {(class, j): (kappa, b)}; returns {(class, j): (info, x)}."""
    B = H.Batch(levels, counts, H.batch_config(**cfg))
    keep = []
    for (c, j), (kappa, b) in systems.items():
        dk, db = H.DeviceArray.from_host(kappa, np.float64), H.DeviceArray.from_host(b)
        keep += [dk, db]
        B.load(B.slot(c, j), dk, db)
    launches = B.run()
    info = B.poll()
    out = {}
    for (c, j), (kappa, b) in systems.items():
        g = B.slot(c, j)
        x = H.DeviceArray(len(b), np.float64)
        B.result(g, x)
        i = info[g]
        out[(c, j)] = ((i.state, i.iters, i.status, i.rel, i.level), x.get())
    return out, launches, B


def test_first_launch_from_history():
    """first_launch_from_history = 1: a batch's first launch runs the last
    batch's largest outer iteration count + 2 iterations.  After a batch whose
    one system has b = 0 (retired at its first check: 2 iterations), the next
    batch needs more than its first launch; the batch after that, the same
    systems again, finishes in its first launch.  Every slot's iterations,
    defect and x equal those of launches of iters_per_launch."""
    levels, counts = [7, 5], [2, 3]
    systems = {(c, j): _system(levels[c], 520 + 10 * c + j) for c in range(2) for j in range(counts[c])}
    ref, nref, _ = _run_batch(levels, counts, systems, iters_per_launch=24)
    assert nref == 1
    B = H.Batch(levels, counts, H.batch_config(iters_per_launch=24, first_launch_from_history=1))
    keep = []

    def load(items):
        for (c, j), (kappa, b) in items.items():
            dk, db = H.DeviceArray.from_host(kappa, np.float64), H.DeviceArray.from_host(b)
            keep.extend([dk, db])
            B.load(B.slot(c, j), dk, db)

    kappa0, b0 = systems[(1, 0)]
    load({(1, 0): (kappa0, np.zeros_like(b0))})
    assert B.run() == 1
    assert B.poll()[B.slot(1, 0)].iters == 0
    for expect_one in (False, True):
        load(systems)
        n = B.run()
        assert (n == 1) if expect_one else (n >= 2), n
        info = B.poll()
        for (c, j) in systems:
            g = B.slot(c, j)
            x = H.DeviceArray(len(systems[(c, j)][1]), np.float64)
            B.result(g, x)
            i = info[g]
            assert (i.state, i.iters, i.status, i.rel, i.level) == ref[(c, j)][0], (expect_one, c, j)
            assert np.array_equal(x.get(), ref[(c, j)][1]), (expect_one, c, j)


def test_fill_fields_match_host_generators():
    n = 10007
    u = H.DeviceArray(n, np.float64)
    H.fill_uniform(u, -1.0, 1.0, 1234)
    assert np.array_equal(u.get(), uniform_pm1(1234, n))
    k = H.DeviceArray(n, np.float64)
    H.fill_lognormal(k, 1.0, 99)
    np.testing.assert_allclose(k.get(), lognormal_cells(99, n), rtol=1e-13)


def test_batch_bitexact_with_single_system_solves():
    """Autotune off (T = 1, block 256), defect-correction outer loop: each
    slot equals the single-system MIXED solve bit for bit -- outer
    iterations, relative defect, x."""
    levels, counts = [3, 5, 4], [2, 2, 3]
    systems = {(c, j): _system(levels[c], 10 * c + j) for c in range(3) for j in range(counts[c])}
    res, launches, B = _run_batch(levels, counts, systems, autotune=0, iters_per_launch=3, outer_method=0)
    assert B.shape(5) == (1, 256, 1)
    for (c, j), ((state, iters, status, rel, level), x) in res.items():
        assert (state, status, level) == (0, 0, levels[c])
        outer, rd, xs = _single(levels[c], *systems[(c, j)])
        assert iters == outer and rel == rd, (c, j, iters, outer, rel, rd)
        assert np.array_equal(x, xs), (c, j, np.abs(x - xs).max())
    assert launches == -(-max(r[0][1] + 1 for r in res.values()) // 3)


def test_batch_residual_fp64_against_oracle():
    """Every slot's FP64 residual, recomputed on the host, is < 1e-8 relative."""
    levels, counts = [6, 4], [2, 2]
    systems = {(c, j): _system(levels[c], 100 + 10 * c + j) for c in range(2) for j in range(counts[c])}
    res, _, _ = _run_batch(levels, counts, systems)
    for (c, j), ((state, iters, status, rel, level), x) in res.items():
        side = 2 ** levels[c] + 1
        kappa, b = systems[(c, j)]
        A = _fp64_operator(side, kappa)
        guess = np.zeros_like(b)
        bd = np.zeros(side * side, bool)
        bd[:side] = bd[-side:] = True
        bd[::side] = True
        guess[bd] = b[bd]
        r0 = np.linalg.norm(oracle.defect(A, b, guess))
        r = np.linalg.norm(oracle.defect(A, b, x))
        assert status == 0 and rel < TOL
        assert r / r0 < 1.001 * TOL, (c, j, r / r0)


def test_batch_alone_equals_in_the_mix():
    """Continuous batching invariant: a system's result is the same solved
    alone and solved in a mixed batch (autotuned launch shapes)."""
    target = _system(5, 7)
    alone, _, _ = _run_batch([5], [1], {(0, 0): target})
    others = {(c, j): _system(lv, 200 + 10 * c + j) for c, lv in enumerate([6, 5, 3]) for j in range(2)}
    others[(1, 1)] = target
    mix, _, _ = _run_batch([6, 5, 3], [2, 2, 2], others)
    (a_info, a_x), (m_info, m_x) = alone[(0, 0)], mix[(1, 1)]
    assert a_info == m_info
    assert np.array_equal(a_x, m_x)


@pytest.mark.parametrize("outer", [0, 1])
def test_fused_coarse_levels_equal_level_by_level(outer):
    """The coarse end of the V-cycle in one launch (levels 1..k, one
    work-group per slot) gives the bits of the level-by-level launches, for
    slots whose finest level is above, at and below k."""
    systems = {(c, j): _system(lv, 300 + 10 * c + j) for c, lv in enumerate([6, 4, 3, 2]) for j in range(2)}
    base, _, _ = _run_batch([6, 4, 3, 2], [2, 2, 2, 2], systems, outer_method=outer, fused_levels=0)
    for k in (2, 4, 6):
        fused, _, _ = _run_batch([6, 4, 3, 2], [2, 2, 2, 2], systems, outer_method=outer, fused_levels=k)
        for key in systems:
            assert base[key][0] == fused[key][0], (k, key)
            assert np.array_equal(base[key][1], fused[key][1]), (k, key)


@pytest.mark.parametrize("outer", [0, 1])
def test_direct_enqueue_equals_graph_launch(outer):
    """use_graphs = 0 enqueues the captured iteration's kernels at every step
    instead of launching the graph: the same bits, iteration counts and
    number of steps."""
    systems = {(c, j): _system(lv, 500 + 10 * c + j) for c, lv in enumerate([6, 4, 2]) for j in range(2)}
    g, lg, _ = _run_batch([6, 4, 2], [2, 2, 2], systems, outer_method=outer, iters_per_launch=3)
    d, ld, _ = _run_batch([6, 4, 2], [2, 2, 2], systems, outer_method=outer, iters_per_launch=3, use_graphs=0)
    assert lg == ld
    for key in systems:
        assert g[key][0] == d[key][0], key
        assert np.array_equal(g[key][1], d[key][1]), key


@pytest.mark.parametrize("inner", [H.F32, H.BF16, H.F64])
@pytest.mark.parametrize("outer", [0, 1])
def test_paired_passes_equal_separate_launches(outer, inner):
    """Levels of 33^2 nodes and up run Jacobi0 + defect and the two
    post-smoothing sweeps as one launch each (pair_sweeps = 1, the default):
    every slot's iterations, defect and x equal those of one launch per pass
    (pair_sweeps = 0), with T = 1 on every level (autotune off)."""
    systems = {(c, j): _system(lv, 700 + 10 * c + j) for c, lv in enumerate([8, 7, 6, 3]) for j in range(2)}
    kw = dict(outer_method=outer, inner_vtype=inner, autotune=0, iters_per_launch=2)
    sep, _, _ = _run_batch([8, 7, 6, 3], [2, 2, 2, 2], systems, pair_sweeps=0, **kw)
    par, _, _ = _run_batch([8, 7, 6, 3], [2, 2, 2, 2], systems, pair_sweeps=1, **kw)
    for key in systems:
        assert sep[key][0] == par[key][0], key
        assert np.array_equal(sep[key][1], par[key][1]), key
        if outer == 1:
            assert sep[key][0][2] == 0, key   # converged


@pytest.mark.parametrize("inner", [H.F32, H.BF16, H.F64])
def test_paired_passes_equal_separate_launches_513_and_1025(inner):
    """The tile shapes of the headline batch's finest level (513^2: 9 column
    strips of 57, 18 row bands of 30 with FP32 / BF16 vectors, 37 of 14 with
    FP64 ones) and of the serving stream's largest size (1025^2: 17 strips of
    61, 35 or 74 bands), one slot each: paired launches and one launch per
    pass give the same iterations, defect and x."""
    systems = {(0, 0): _system(10, 900), (1, 0): _system(9, 901)}
    kw = dict(outer_method=1, inner_vtype=inner, autotune=0, iters_per_launch=4)
    sep, _, _ = _run_batch([10, 9], [1, 1], systems, pair_sweeps=0, **kw)
    par, _, _ = _run_batch([10, 9], [1, 1], systems, pair_sweeps=1, **kw)
    for key in systems:
        assert sep[key][0] == par[key][0], key
        assert sep[key][0][2] == 0, key
        assert np.array_equal(sep[key][1], par[key][1]), key


@pytest.mark.parametrize("inner", [H.F32, H.BF16, H.F64])
@pytest.mark.parametrize("outer", [0, 1])
def test_matrix_free_equals_stored_values(outer, inner):
    """matrix_free = 1 (default): the paired launches and the FP64 outer row
    sums form each row's weights from the slots' cell coefficients with the
    assembly's expressions; matrix_free = 2: the paired launches gather them
    from the per-node arrays of distinct weights.  Every slot's iterations,
    defect and x equal those of matrix_free = 0 (values read from the stored
    matrices), bit for bit, at sides 9..513."""
    levels = [9, 7, 5, 3]
    systems = {(c, j): _system(lv, 800 + 10 * c + j) for c, lv in enumerate(levels) for j in range(2)}
    kw = dict(outer_method=outer, inner_vtype=inner, iters_per_launch=3)
    st, _, _ = _run_batch(levels, [2] * 4, systems, matrix_free=0, **kw)
    for form in (1, 2):
        mf, _, _ = _run_batch(levels, [2] * 4, systems, matrix_free=form, **kw)
        for key in systems:
            assert st[key][0] == mf[key][0], (form, key)
            assert np.array_equal(st[key][1], mf[key][1]), (form, key)
            if outer == 1:
                assert mf[key][0][2] == 0, (form, key)


def test_batch_slot_reuse_and_late_admission():
    """A slot reloaded after its first system retired, while another slot is
    still running, solves the new system exactly as a fresh batch does."""
    sys_a, sys_b, sys_c = _system(4, 1), _system(4, 2), _system(6, 3)
    B = H.Batch([4, 6], [1, 1], H.batch_config(iters_per_launch=1))
    ga, gc = B.slot(0, 0), B.slot(1, 0)
    keep = [H.DeviceArray.from_host(a, np.float64) for a in (*sys_a, *sys_c)]
    B.load(ga, keep[0], keep[1])
    B.load(gc, keep[2], keep[3])
    while True:   # until slot a has been admitted and retired
        B.step()
        info = B.poll()
        if info[ga].status != -1:
            break
    assert info[gc].state != 0, "the level-6 system should still be running"
    xa = H.DeviceArray(17 * 17, np.float64)
    B.result(ga, xa)
    kb = [H.DeviceArray.from_host(a, np.float64) for a in sys_b]
    B.load(ga, kb[0], kb[1])
    B.run()
    info = B.poll()
    xb = H.DeviceArray(17 * 17, np.float64)
    B.result(ga, xb)
    fresh_a, _, _ = _run_batch([4], [1], {(0, 0): sys_a}, iters_per_launch=1)
    fresh_b, _, _ = _run_batch([4], [1], {(0, 0): sys_b}, iters_per_launch=1)
    assert np.array_equal(xa.get(), fresh_a[(0, 0)][1])
    assert np.array_equal(xb.get(), fresh_b[(0, 0)][1])
    assert (info[ga].iters, info[ga].rel) == fresh_b[(0, 0)][0][1:4:2]


def test_batch_not_converged_and_budget():
    res, _, _ = _run_batch([5], [1], {(0, 0): _system(5, 9)}, max_outer=2, iters_per_launch=4)
    (state, iters, status, rel, level), _ = res[(0, 0)]
    assert (state, iters, status) == (0, 2, 5) and rel > TOL
    with pytest.raises(H.HipError) as e:
        H.Batch([9], [4], H.batch_config(), budget_bytes=1 << 20)
    assert e.value.code == 6
    with pytest.raises(H.HipError) as e:
        H.Batch([1], [1], H.batch_config())
    assert e.value.code == 1
    for bad in (dict(fused_levels=7), dict(fused_levels=-1), dict(outer_method=2), dict(n_pre=1),
                dict(iters_per_launch=0), dict(averaging=3), dict(use_graphs=2),
                dict(pair_sweeps=2), dict(matrix_free=3), dict(matrix_free=-1), dict(inner_vtype=9)):
        with pytest.raises(H.HipError) as e:
            H.Batch([5], [1], H.batch_config(**bad))
        assert e.value.code == 1, bad
    with pytest.raises(H.HipError) as e:   # finest level above the pool's limit
        H.Batch([14], [1], H.batch_config())
    assert e.value.code == 1


def test_batch_bf16_inner_converges():
    """BF16 matrix values in the V-cycle (FP32 vectors and arithmetic): the
    FP64 outer loop still reaches 1e-8."""
    levels, counts = [6, 5], [1, 2]
    systems = {(c, j): _system(levels[c], 300 + 10 * c + j) for c in range(2) for j in range(counts[c])}
    res, _, _ = _run_batch(levels, counts, systems, inner_vtype=H.BF16)
    f32, _, _ = _run_batch(levels, counts, systems)
    for key, ((state, iters, status, rel, level), x) in res.items():
        assert status == 0 and rel < TOL
        x32 = f32[key][1]
        assert np.abs(x - x32).max() < 1e-5 * max(1.0, np.abs(x32).max())


def test_serve_same_alone_and_in_the_mix():
    """Continuous batching: a request's solution is bit-identical when served
    alone and when it arrives while other requests of several sizes are
    already running (and slots are reused)."""
    levels = [4, 5, 6]
    target = (0.0, 5, 777, 778)
    xa = H.DeviceArray(33 * 33, np.float64)
    res, st = H.serve(levels, [target + (xa,)], counts=[2, 2, 2])
    assert res[0].status == 0 and st.completed == 1
    mix = []
    for i in range(12):
        lv = levels[i % 3]
        mix.append((0.002 * i, lv, 1000 + 2 * i, 1001 + 2 * i))
    xm = H.DeviceArray(33 * 33, np.float64)
    mix.insert(7, (0.002 * 6.5, 5, 777, 778, xm))
    res, st = H.serve(levels, mix, counts=[2, 2, 2])
    assert st.completed == len(mix) and st.rejected == 0 and st.not_converged == 0
    for r, q in zip(res, mix):
        assert r.status == 0 and r.rel < TOL
        assert q[0] <= r.admit_s <= r.done_s
    assert np.array_equal(xa.get(), xm.get())
    assert st.p50_ms <= st.p99_ms <= st.max_ms


def test_serve_budget_and_unknown_size():
    with pytest.raises(H.HipError) as e:
        H.serve([9], [(0.0, 9, 1, 2)], counts=[8], budget_bytes=64 << 20)
    assert e.value.code == 6
    res, st = H.serve([4], [(0.0, 4, 1, 2), (0.0, 7, 3, 4), (0.001, 4, 5, 6)], budget_bytes=1 << 30,
                      max_slots_per_class=1)
    assert [r.status for r in res] == [0, -2, 0]
    assert st.completed == 2 and st.rejected == 1
    assert res[2].admit_s >= res[0].done_s - 1e-9   # one slot: the second level-4 request waits for it


@pytest.mark.parametrize("inner", [H.F32, H.F64])
def test_batch_cg_matches_single_system_cg(inner):
    """Batched CG outer loop (default) against the single-system CG driver on
    the same hierarchy (FP32 inner, or everything in FP64): same iteration
    count within one, solutions agree to the accuracy the tolerance implies
    (1e-5 relative with the FP32 inner cycle, 1e-7 in FP64)."""
    levels, counts = [6, 5], [1, 2]
    systems = {(c, j): _system(levels[c], 400 + 10 * c + j) for c in range(2) for j in range(counts[c])}
    res, _, _ = _run_batch(levels, counts, systems, autotune=0, inner_vtype=inner)
    for (c, j), ((state, iters, status, rel, level), x) in res.items():
        assert status == 0 and rel < TOL
        L = levels[c]
        n = (2 ** L + 1) ** 2
        kappa, b = systems[(c, j)]
        mg = H.Multigrid(inner, 1, L)
        mg.configure(H.mg_config(tolerance=TOL, max_outer=200))
        mg.build_q1_coeff(H.DeviceArray.from_host(kappa, np.float64), OMEGA, AVG)
        db, dx = H.DeviceArray.from_host(b), H.DeviceArray(n, np.float64)
        dx.zero()
        it, rd = mg.solve_pcg(db, dx)
        assert abs(it - iters) <= 1, (c, j, it, iters)
        xs = dx.get()
        tol = 1e-5 if inner == H.F32 else 1e-7
        assert np.abs(x - xs).max() <= tol * max(1.0, np.abs(xs).max())


def test_batch_cg_converges_where_defect_correction_diverges():
    """System 48 of the 256-system lognormal workload (seed 4, side 513): the
    rediscretised hierarchy over-corrects and plain defect correction
    diverges; the CG outer loop converges (FP64 residual recomputed)."""
    side, n = 513, 513 * 513
    cs, bs = poisson_system_seeds(4, 48)
    k = H.DeviceArray((side - 1) ** 2, np.float64)
    b = H.DeviceArray(n, np.float64)
    H.fill_lognormal(k, 1.0, cs)
    H.fill_uniform(b, -1.0, 1.0, bs)
    H.synchronize()
    out = {}
    for method in (0, 1):
        B = H.Batch([9], [1], H.batch_config(outer_method=method, max_outer=60))
        B.load(0, k, b)
        B.run()
        x = H.DeviceArray(n, np.float64)
        B.result(0, x)
        out[method] = (B.poll()[0], x.get())
    assert out[0][0].status == 5 and out[0][0].rel > 1.0
    info, x = out[1]
    assert info.status == 0 and info.rel < TOL and info.iters < 30
    kappa, bh = k.get(), b.get()
    A = _fp64_operator(side, kappa)
    bd = np.zeros(n, bool)
    bd[:side] = bd[-side:] = True
    bd[::side] = True
    r0 = np.linalg.norm(oracle.defect(A, bh, np.where(bd, bh, 0.0)))
    assert np.linalg.norm(oracle.defect(A, bh, x)) / r0 < 1.001 * TOL


def test_bulk_loads_equal_single_loads():
    """One bulk build of several slots (from device arrays, and from seeds)
    gives the same solutions, bit for bit, as loading the slots one by one
    from the same data."""
    levels, counts = [5, 4], [2, 2]
    seeds = [poisson_system_seeds(21, i) for i in range(4)]
    data = []
    for i, L in enumerate([5, 5, 4, 4]):
        side = 2 ** L + 1
        k = H.DeviceArray((side - 1) ** 2, np.float64)
        b = H.DeviceArray(side * side, np.float64)
        H.fill_lognormal(k, 1.0, seeds[i][0])
        H.fill_uniform(b, -1.0, 1.0, seeds[i][1])
        data.append((k, b))
    H.synchronize()
    outs = []
    for mode in ("single", "many", "seeded"):
        B = H.Batch(levels, counts, H.batch_config())
        slots = [B.slot(0, 0), B.slot(0, 1), B.slot(1, 0), B.slot(1, 1)]
        if mode == "single":
            for g, (k, b) in zip(slots, data):
                B.load(g, k, b)
        elif mode == "many":
            B.load_many(slots, [k for k, _ in data], [b for _, b in data])
        else:
            B.load_seeded(slots, [s[0] for s in seeds], [s[1] for s in seeds])
        B.run()
        info = B.poll()
        xs = []
        for g, (k, b) in zip(slots, data):
            x = H.DeviceArray(b.n, np.float64)
            B.result(g, x)
            xs.append((info[g].iters, info[g].status, x.get()))
        outs.append(xs)
    for other in outs[1:]:
        for (i0, s0, x0), (i1, s1, x1) in zip(outs[0], other):
            assert (i0, s0) == (i1, s1) and s0 == 0
            assert np.array_equal(x0, x1)
