"""Full-size parity of the batched (256 systems, sides 33..513) and served
(Poisson stream, sides 65..1025) workloads.

Each system is one mixed-precision multigrid solve of a lognormal-coefficient
Q1 problem; per system the numerics are the reference's MIXED defect
correction (math/multigrid.hh:2395-2536) or its CG-outer variant, so the bars
are those of the reference's solver tests (math/mg_TEST.cc:389-464 compares
each result with a reference solution):
  * every result's FP64 relative residual ||b - A x|| / ||b - A x0||,
    recomputed on the host by the oracle's Defect<tags::CPU> on the oracle's
    own assembly of the operator from the coefficient the device drew, is
    below the tolerance (x0: the boundary start vector of the driver);
  * the device's FP64 operator is the oracle's, bit for bit;
  * the error against the oracle's FP64 direct solve (sparse LU plus
    refinement, aiming at a relative residual of 1e-14 and reaching at most
    1e-13 -- the rounding of the FP64 defect itself -- five orders below the
    solver tolerance) is stated for a sample and bounded;
  * a request's result is bit-identical served alone and served inside the
    stream, at 257^2 and 1025^2.
"""
import numpy as np
import pytest

import honei_amd as H
import oracle
from tests.patterns import SERVE_LEVELS, batch_levels, poisson_system_seeds, serving_requests

pytestmark = pytest.mark.gpu

TOL = 1e-8
# Error of a 1e-8-residual solution against the FP64 direct solve, relative
# to max |x64|: bounded by the solution's residual times the conditioning of
# the lognormal operators; the measured values are printed by the tests.
ERR_BOUND = 1e-6


@pytest.fixture(scope="module", autouse=True)
def _init():
    H.init(0)
    yield


def _device_fields(level, cs, bs):
    """The coefficient and right-hand side the device draws from the seeds:
    the fields the seeded bulk loads and the serving loop assemble."""
    side = 2 ** level + 1
    k = H.DeviceArray((side - 1) ** 2, np.float64)
    b = H.DeviceArray(side * side, np.float64)
    H.fill_lognormal(k, 1.0, cs)
    H.fill_uniform(b, -1.0, 1.0, bs)
    H.synchronize()
    return k.get(), b.get()


def _boundary_guess(side, b):
    """Start vector of the MIXED driver (math/multigrid.hh:2423-2440): b on
    the first and last grid rows and the first grid column (the last column
    stays 0, as the reference leaves it)."""
    g = np.zeros_like(b)
    bd = np.zeros(side * side, bool)
    bd[:side] = bd[-side:] = True
    bd[::side] = True
    g[bd] = b[bd]
    return g


def _fp64_check(level, kappa, b, x):
    """(oracle FP64 relative residual of x, the oracle's operator)."""
    side = 2 ** level + 1
    A = oracle.q1_coeff_matrix(side, kappa)
    r0 = np.linalg.norm(oracle.defect(A, b, _boundary_guess(side, b)))
    return float(np.linalg.norm(oracle.defect(A, b, x)) / r0), A


def _err_vs_direct(A, b, x):
    x64, rel, _ = oracle.direct_solve(A, b)
    assert rel <= 1e-13
    return float(np.abs(x - x64).max() / np.abs(x64).max())


@pytest.mark.parametrize("level", list(SERVE_LEVELS))
def test_device_fp64_operator_is_the_oracles(level):
    """The FP64 outer operator the batch assembles from a lognormal
    coefficient equals the oracle's assembly bit for bit (index and value
    arrays), at every serving size."""
    side = 2 ** level + 1
    n = side * side
    cs, bs = poisson_system_seeds(5, level)
    kappa, _ = _device_fields(level, cs, bs)
    stride = 32 * ((n + 31) // 32)
    Aj = H.DeviceArray(9 * stride, np.int32)
    Ax = H.DeviceArray(9 * stride, np.float64)
    Arl = H.DeviceArray(n, np.uint32)
    dk = H.DeviceArray.from_host(kappa, np.float64)
    H.call("honei_hip_q1_assemble", H.F64, side, Aj.ptr, Ax.ptr, Arl.ptr, stride, dk.ptr, None, 0.0, 1, None)
    A = oracle.q1_coeff_matrix(side, kappa)
    assert A.stride == stride and A.ncpr == 9
    assert np.array_equal(Aj.get().astype(np.uint64), A.Aj)
    assert np.array_equal(Arl.get().astype(np.uint64), A.Arl)
    assert np.array_equal(Ax.get().view(np.uint64), A.Ax.view(np.uint64))


def _serve_with_solutions(reqs, counts, **kw):
    xs = [H.DeviceArray((2 ** r[1] + 1) ** 2, np.float64) for r in reqs]
    res, st = H.serve(list(SERVE_LEVELS), [r + (x,) for r, x in zip(reqs, xs)], counts=counts, **kw)
    return res, st, xs


def test_serve_stream_sides_65_to_1025_checked_by_oracle():
    """The first 40 requests of the serving stream (Poisson arrivals at
    2000/s, sides 65..1025 including 1025^2, seed 5) through a pool of two
    slots per size, so requests queue, overlap on the device and reuse
    slots: all converge, and the oracle's FP64 residual of every served
    solution is below 1e-8."""
    reqs = serving_requests(1.0)[:40]
    assert {r[1] for r in reqs} == set(SERVE_LEVELS)
    res, st, xs = _serve_with_solutions(reqs, [2] * 5)
    print(f"served {st.completed} in {st.wall_s:.3f} s, {st.launches} launches, max queue {st.max_queue}", flush=True)
    assert st.completed == len(reqs) and st.not_converged == 0 and st.rejected == 0
    assert st.max_queue > 0
    worst = {}
    for r, q, x in zip(res, reqs, xs):
        assert r.status == 0 and r.rel < TOL, (q, r.status, r.rel)
        assert q[0] <= r.admit_s <= r.done_s
        kappa, b = _device_fields(q[1], q[2], q[3])
        rel, _ = _fp64_check(q[1], kappa, b, x.get())
        worst[q[1]] = max(worst.get(q[1], 0.0), rel)
        assert rel < 1.001 * TOL, (q, rel, r.rel)
    print("oracle FP64 residual, worst per level:", {k: f"{v:.2e}" for k, v in sorted(worst.items())}, flush=True)


def test_serve_alone_equals_in_the_stream_257_and_1025():
    """Continuous batching at the stream's sizes: a 257^2 and a 1025^2
    request give the same bits served alone (one slot per size) and served
    while the stream's other requests are on the device (two slots per
    size, queueing, slot reuse)."""
    targets = []
    for k, level in enumerate((8, 10)):
        cs, bs = poisson_system_seeds(55, k)
        targets.append((level, cs, bs))
    alone = []
    for level, cs, bs in targets:
        res, st, xs = _serve_with_solutions([(0.0, level, cs, bs)], [1] * 5)
        assert res[0].status == 0 and st.completed == 1
        alone.append((res[0].iters, res[0].rel, xs[0].get()))
    stream = serving_requests(1.0)[:24]
    mix = list(stream)
    mix.insert(9, (stream[8][0], targets[0][0], targets[0][1], targets[0][2]))
    mix.insert(17, (stream[15][0], targets[1][0], targets[1][1], targets[1][2]))
    res, st, xs = _serve_with_solutions(mix, [2] * 5)
    assert st.completed == len(mix) and st.not_converged == 0
    for pos, (it, rel, xa) in zip((9, 17), alone):
        r = res[pos]
        assert r.status == 0 and r.admit_s >= mix[pos][0]
        assert (r.iters, r.rel) == (it, rel), (pos, r.iters, it, r.rel, rel)
        assert np.array_equal(xs[pos].get(), xa), pos


def test_serve_host_payload_equals_seeded():
    """Requests whose coefficient and right-hand side come from pinned host
    memory and whose solution returns there, mixed with seeded requests in
    one stream (sides 65..1025): every result has the bits, iterations and
    residual of the same system served from its seeds; a payload request is
    done only after its solution has reached host memory."""
    stream = serving_requests(1.0)[:20]
    res0, st0, xs0 = _serve_with_solutions(stream, [2] * 5)
    assert st0.completed == len(stream)
    mix, host = [], {}
    for i, q in enumerate(stream):
        if i % 2:
            kappa, b = _device_fields(q[1], q[2], q[3])
            pk, pb, px = H.PinnedArray(kappa.size), H.PinnedArray(b.size), H.PinnedArray(b.size)
            pk.a[:] = kappa
            pb.a[:] = b
            px.a[:] = np.nan
            host[i] = px
            mix.append(q + (None, pk, pb, px))
        else:
            host[i] = H.DeviceArray(xs0[i].n, np.float64)
            mix.append(q + (host[i],))
    res, st = H.serve(list(SERVE_LEVELS), mix, counts=[2] * 5)
    assert st.completed == len(mix) and st.not_converged == 0
    for i, (r, r0) in enumerate(zip(res, res0)):
        x = host[i].a.copy() if i % 2 else host[i].get()
        assert (r.iters, r.rel, r.status) == (r0.iters, r0.rel, r0.status), i
        assert np.array_equal(x, xs0[i].get()), i
        assert r.admit_s <= r.done_s


@pytest.mark.parametrize("levels", [(6, 7, 8, 9), (10,)])
def test_serve_error_vs_fp64_direct_solve(levels):
    """One request per size of the stream: the served solution against the
    oracle's FP64 direct solve of the same system (final error of the
    result, as math/mg_TEST.cc:389-464 compares with a reference
    solution)."""
    stream = serving_requests(1.0)
    reqs = [next(q for q in stream if q[1] == L) for L in levels]
    reqs = [(0.001 * i,) + q[1:] for i, q in enumerate(reqs)]
    res, st, xs = _serve_with_solutions(reqs, [1] * 5)
    assert st.completed == len(reqs)
    for r, q, x in zip(res, reqs, xs):
        kappa, b = _device_fields(q[1], q[2], q[3])
        rel, A = _fp64_check(q[1], kappa, b, x.get())
        err = _err_vs_direct(A, b, x.get())
        print(f"level {q[1]}: device rel {r.rel:.3e}, oracle rel {rel:.3e}, error vs FP64 direct {err:.3e}",
              flush=True)
        assert rel < 1.001 * TOL and err < ERR_BOUND, (q[1], rel, err)


def _config4_pool(inner):
    levels = [int(v) for v in batch_levels(256, 4)]
    lv = sorted(set(levels), reverse=True)
    B = H.Batch(lv, [levels.count(v) for v in lv], H.batch_config(inner_vtype=inner))
    nxt = {v: 0 for v in lv}
    slots = []
    for L in levels:
        slots.append(B.slot(lv.index(L), nxt[L]))
        nxt[L] += 1
    seeds = [poisson_system_seeds(4, i) for i in range(256)]
    B.load_seeded(slots, [s[0] for s in seeds], [s[1] for s in seeds])
    launches = B.run()
    return B, levels, slots, seeds, launches


@pytest.mark.parametrize("inner", [H.F32, H.BF16])
def test_batch_256_systems_checked_by_oracle(inner):
    """The whole batched workload (256 systems, sides 33..513 log-uniform,
    lognormal coefficient, seed 4) in one pool: every slot converges and the
    oracle's FP64 residual of every solution is below 1e-8; the FP64 direct
    solve bounds the error of two systems per size."""
    B, levels, slots, seeds, launches = _config4_pool(inner)
    info = B.poll()
    print(f"inner {inner}: {launches} launches", flush=True)
    worst, errs = {}, {}
    for i, (L, g, (cs, bs)) in enumerate(zip(levels, slots, seeds)):
        assert info[g].status == 0 and info[g].rel < TOL and info[g].level == L, (i, info[g].status, info[g].rel)
        x = H.DeviceArray((2 ** L + 1) ** 2, np.float64)
        B.result(g, x)
        kappa, b = _device_fields(L, cs, bs)
        xh = x.get()
        rel, A = _fp64_check(L, kappa, b, xh)
        worst[L] = max(worst.get(L, 0.0), rel)
        assert rel < 1.001 * TOL, (i, L, rel, info[g].rel)
        if len(errs.setdefault(L, [])) < 2:
            errs[L].append(_err_vs_direct(A, b, xh))
    print("oracle FP64 residual, worst per level:", {k: f"{v:.2e}" for k, v in sorted(worst.items())}, flush=True)
    print("error vs FP64 direct solve:", {k: [f"{e:.2e}" for e in v] for k, v in sorted(errs.items())}, flush=True)
    assert max(max(v) for v in errs.values()) < ERR_BOUND


def test_batch_destroy_right_after_step():
    """A pool destroyed while its last launch may still be running (no poll
    in between) waits for that work before releasing the graphs and the
    memory; a new pool then solves normally."""
    for _ in range(2):
        B = H.Batch([9, 6], [2, 2], H.batch_config())
        seeds = [poisson_system_seeds(66, i) for i in range(4)]
        B.load_seeded([B.slot(0, 0), B.slot(0, 1), B.slot(1, 0), B.slot(1, 1)], [s[0] for s in seeds],
                      [s[1] for s in seeds])
        B.activate(True)
        B.step()
        B.close()
    B = H.Batch([6], [1], H.batch_config())
    cs, bs = poisson_system_seeds(66, 9)
    B.load_seeded([0], [cs], [bs])
    B.run()
    assert B.poll()[0].status == 0
