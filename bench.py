#!/usr/bin/env python3
"""Benchmark of the MI355X backend for HONEI's ELL sparse-solve path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--serve-seconds S]

One step = one batched solve: 256 independent Q1 Poisson systems, side
2^k+1 with k uniform in {5..9} (33..513 nodes), a lognormal (sigma 1)
per-cell diffusion coefficient and a right-hand side uniform in [-1, 1)
(synthetic, seeded, drawn on the device), each solved to a relative FP64
residual of 1e-8 by mixed-precision multigrid: FP64 conjugate gradients
(defect, update, norms in FP64) preconditioned by one FP32 V(2,2)
damped-Jacobi cycle per iteration, in graph launches over a
device-resident slot pool.  The timed step covers the
assembly of all 256 hierarchies from the resident coefficients and the
solve.  `value` is solves/s; p50/p99 of the per-system completion time
within a step are reported beside it.

The same run measures: the plain single-system FP64 GPU multigrid solve
of the same 256 systems (one at a time), the continuous-batching serving
loop (Poisson arrivals at 2000 requests/s, sides 65..1025, 64 GB pool
budget; p50/p99 latency), the single 2049^2 mixed solve, the standalone
ELL SpMV (Q1 2049^2 and a random 4M-row ELL, FP32/FP64) as a fraction of
HBM peak (`roofline`), and the CPU path on the host (`cpu_baseline`).

Prints exactly one JSON line on stdout (progress goes to stderr).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import honei_amd as H  # noqa: E402
from tests import patterns  # noqa: E402
from tests.patterns import batch_levels, poisson_system_seeds, uniform01, uniform_pm1  # noqa: E402

METRIC = "Poisson solves/s + p99 latency at 1e-8 rel. residual; ELL SpMV % of HBM peak"
PEAK_GBS = 8000.0
SIDE = 2049
LEVEL = 11
NSYS, BATCH_SEED = 256, 4          # batched workload
SERVE_RATE, SERVE_SEED = 2000.0, 5  # serving workload
SERVE_FP64_SAMPLE = 250  # requests of the stream timed with the plain FP64 solve
SERVE_LEVELS = [6, 7, 8, 9, 10]     # sides 65 .. 1025
SERVE_BUDGET = 64 << 30
ERR_SAMPLE = 2                      # systems per size whose error vs the FP64 direct solve is stated


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_q1(side, vt, with_dinv=False):
    """Q1 operator assembled on the device (bit-identical to the reference's
    FillMatrix -> SparseMatrix -> SparseMatrixELL chain); with_dinv also
    fills D.dinv = 1 / diag(A)."""
    n = side * side
    stride = 32 * ((n + 31) // 32)

    class E:
        pass

    D = E()
    D.rows, D.cols, D.stride, D.ncpr, D.threads, D.vtype = n, n, stride, 9, 1, vt
    D.Aj = H.DeviceArray(9 * stride, np.int32)
    D.Ax = H.DeviceArray(9 * stride, np.float64 if vt == H.F64 else np.float32)
    D.Arl = H.DeviceArray(n, np.uint32)
    D.dinv = H.DeviceArray(n, np.float64 if vt == H.F64 else np.float32) if with_dinv else None
    H.call("honei_hip_q1_assemble", vt, side, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, stride, None,
           D.dinv.ptr if with_dinv else None, 1.0, 1, None)
    D.nnz = 9 * (side - 2) ** 2 + 4 * (side - 1)
    return D


def device_random_ell(n, vt, seed=2, max_len=27, band=4096, threads=1):
    """Random ELL of the SpMV configuration generated in HBM (rows U{1..27}
    long, distinct columns within +-4096 of the diagonal, N(0,1)), in the
    T-interleaved ELL layout of `threads`."""
    stride = 32 * ((n * threads + 31) // 32)
    ncpr = (max_len + threads - 1) // threads

    class E:
        pass

    D = E()
    D.rows, D.cols, D.stride, D.ncpr, D.threads, D.vtype = n, n, stride, ncpr, threads, vt
    D.Aj = H.DeviceArray(ncpr * stride, np.int32)
    D.Ax = H.DeviceArray(ncpr * stride, np.float64 if vt == H.F64 else np.float32)
    D.Arl = H.DeviceArray(n, np.uint32)
    H.call("honei_hip_random_ell", vt, n, max_len, band, seed, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, stride, threads, None)
    D.nnz = int(np.count_nonzero(D.Ax.get()))
    return D


def time_spmv(sets, reps, label, idx64=False):
    """Average duration of one ELL SpMV launch (HIP events on the launch
    stream) and algorithmic bytes per launch nnz*(s_v+s_i) + n*(s_x+s_y),
    s_i = 4 (8 with idx64: the reference's unsigned long Aj read by the
    *_idx64 entry).  `sets` holds (matrix, x, y) operand copies; the
    launches alternate between them, so no launch finds its operands in the
    256 MB MALL from the launch before."""
    D0 = sets[0][0]
    sfx = "float" if D0.vtype == H.F32 else "double"
    fn = f"honei_hip_product_smell_dv_{sfx}" + ("_idx64" if idx64 else "")

    def args(k):
        D, x, y = sets[k % len(sets)]
        Aj, Arl = (D.Aj64, D.Arl64) if idx64 else (D.Aj, D.Arl)
        return (x.ptr, y.ptr, Aj.ptr, D.Ax.ptr, Arl.ptr, 0, D.rows, D.ncpr, D.stride, 0, D.threads, None)
    for k in range(2 * len(sets)):
        H.call(fn, *args(k))
    H.synchronize()
    e0, e1 = H.Event(), H.Event()
    e0.record()
    for k in range(reps):
        H.call(fn, *args(k))
    e1.record()
    ms = e0.elapsed_ms(e1) / reps
    sv = 8 if D0.vtype == H.F64 else 4
    nnz = sum(st[0].nnz for st in sets) / len(sets)   # the copies may differ in pattern
    alg = nnz * (sv + (8 if idx64 else 4)) + D0.rows * 2 * sv
    gbs = alg / (ms * 1e-3) / 1e9
    log(f"[spmv] {label}: {ms * 1e3:.1f} us/launch, {alg / 1e6:.1f} MB, {gbs:.0f} GB/s = {100 * gbs / PEAK_GBS:.1f}%")
    return {"us_per_launch": round(ms * 1e3, 2), "algorithmic_bytes": int(alg), "achieved_gbs": round(gbs, 1),
            "pct_hbm_peak": round(100 * gbs / PEAK_GBS, 2), "operand_copies": len(sets)}


def time_sell(sets, reps, label, stage_x=True):
    """time_spmv for the sliced copy of each operand matrix
    (honei_hip_product_sell; the copies are built outside the timed
    launches); the same algorithmic bytes as the ELL product, and the slots
    actually stored beside them.  stage_x: each 2048-row window's column
    range of x copied into LDS (the default) or x gathered from global
    memory."""
    sells = [(H.Sell(D), x, y) for D, x, y in sets]
    for S, _, _ in sells:
        S.stage_x(stage_x)
    label = label + ("" if stage_x else ", x from global memory")
    for k in range(2 * len(sells)):
        S, x, y = sells[k % len(sells)]
        S.product(y, x)
    H.synchronize()
    e0, e1 = H.Event(), H.Event()
    e0.record()
    for k in range(reps):
        S, x, y = sells[k % len(sells)]
        S.product(y, x)
    e1.record()
    ms = e0.elapsed_ms(e1) / reps
    D0 = sets[0][0]
    sv = 8 if D0.vtype == H.F64 else 4
    nnz = sum(st[0].nnz for st in sets) / len(sets)
    alg = nnz * (sv + 4) + D0.rows * 2 * sv
    pad = sum(S.padded_slots for S, _, _ in sells) / len(sells)
    gbs = alg / (ms * 1e-3) / 1e9
    log(f"[spmv] {label} sliced: {ms * 1e3:.1f} us/launch, {gbs:.0f} GB/s = {100 * gbs / PEAK_GBS:.1f}% "
        f"({pad / nnz:.3f} stored slots per non-zero)")
    out = {"us_per_launch": round(ms * 1e3, 2), "algorithmic_bytes": int(alg), "achieved_gbs": round(gbs, 1),
           "pct_hbm_peak": round(100 * gbs / PEAK_GBS, 2), "operand_copies": len(sets),
           "stored_slots_per_nonzero": round(pad / nnz, 4),
           "what": ("honei_hip_product_sell: rows sorted by length within 2048-row windows, 64-row slices "
                    "padded to their longest row, terms in stored order (bit-identical to the ELL product); "
                    + ("x gathered from an LDS copy of each window's column range" if stage_x
                       else "x gathered from global memory"))}
    del sells
    return out


def with_idx64(D):
    """Attach the reference's 64-bit index arrays (Aj, Arl as unsigned long)
    of a device matrix, widened on the host from its 32-bit copies."""
    D.Aj64 = H.DeviceArray.from_host(D.Aj.get().view(np.uint32).astype(np.uint64))
    D.Arl64 = H.DeviceArray.from_host(D.Arl.get().astype(np.uint64))
    return D


def hbm_ceiling(reps=20):
    """Measured streaming ceilings over 1 GiB arrays, HIP events on the
    launch stream: the probe kernels (16-byte non-temporal loads, eight in
    flight per thread, honei_hip_stream_probe) reading (bytes = read) and
    copying (bytes = read + write), at two grid sizes, beside hipMemcpy
    device-to-device."""
    nbytes = 1 << 30
    a = H.DeviceArray(nbytes // 4, np.float32)
    b = H.DeviceArray(nbytes // 4, np.float32)
    part = H.DeviceArray(65536, np.float32)   # one sum per work-group (grids up to 65536)
    a.zero()
    b.zero()
    out = {}
    cases = [("d2d_copy_hipmemcpy", lambda: H.call("honei_hip_copy", b.ptr, a.ptr, nbytes, None), 2 * nbytes)]
    for grid in (2048, 8192):
        cases.append((f"probe_read_grid{grid}",
                      lambda g=grid: H.call("honei_hip_stream_probe", 0, part.ptr, a.ptr, nbytes, g, None), nbytes))
        cases.append((f"probe_copy_grid{grid}",
                      lambda g=grid: H.call("honei_hip_stream_probe", 1, b.ptr, a.ptr, nbytes, g, None), 2 * nbytes))
    for grid in (8192, 32768):   # one contiguous chunk per work-group, chunks contiguous per XCD
        cases.append((f"probe_read_chunked{grid}",
                      lambda g=grid: H.call("honei_hip_stream_probe", 2, part.ptr, a.ptr, nbytes, g, None), nbytes))
        cases.append((f"probe_copy_chunked{grid}",
                      lambda g=grid: H.call("honei_hip_stream_probe", 3, b.ptr, a.ptr, nbytes, g, None), 2 * nbytes))
    for label, fn, nb in cases:
        for _ in range(3):
            fn()
        H.synchronize()
        e0, e1 = H.Event(), H.Event()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        ms = e0.elapsed_ms(e1) / reps
        gbs = nb / (ms * 1e-3) / 1e9
        out[label] = {"us": round(ms * 1e3, 1), "gbs": round(gbs, 1), "pct_hbm_peak": round(100 * gbs / PEAK_GBS, 2)}
        log(f"[hbm] {label}: {gbs:.0f} GB/s")
    out["read_ceiling_gbs"] = max(v["gbs"] for k, v in out.items() if k.startswith("probe_read"))
    out["copy_ceiling_gbs"] = max(v["gbs"] for k, v in out.items() if k.startswith("probe_copy"))
    del a, b, part
    return out


def cpu_spmv_q1(side, reps=5):
    """The oracle's C restatement of Product<tags::CPU::Generic> (one core)
    and mc::Product (row blocks over all host cores) on the Q1 side^2 FP32
    operator: median wall time per product."""
    import oracle
    A = oracle.q1_matrix(side, np.float32)
    x = uniform_pm1(2, A.rows).astype(np.float32)
    parts = os.cpu_count() or 1
    res = {}
    for label, fn in (("generic_1_core", lambda: oracle.product(A, x)),
                      ("mc_all_cores", lambda: oracle.product_mc(A, x, parts))):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        ms = 1e3 * float(np.median(ts))
        res[label] = {"ms": round(ms, 3), "gbs": round(335347792 / (ms * 1e-3) / 1e9, 2)}
    res["cores"] = parts
    res["kind"] = "port"
    log(f"[cpu spmv] 1 core {res['generic_1_core']['ms']} ms, {parts} cores {res['mc_all_cores']['ms']} ms")
    return res


def time_spmv_staged(D, reps, x, y, label, block=1024, chunks=1):
    """time_spmv for the product with each work-group's x window staged in
    LDS (column windows computed once per matrix, outside the timed
    launches)."""
    rpb = block // D.threads * chunks
    nb = (D.rows + rpb - 1) // rpb
    ranges = H.DeviceArray(2 * nb, np.int32)
    import ctypes
    w = ctypes.c_int64()
    H.call("honei_hip_ell_block_ranges", D.Aj.ptr, D.Arl.ptr, D.rows, D.ncpr, D.stride, D.threads, rpb, ranges.ptr,
           ctypes.byref(w), None)
    vs = 8 if D.vtype == H.F64 else 4
    window = min(w.value, 98304 // vs)
    args = (D.vtype, x.ptr, y.ptr, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, D.rows, D.ncpr, D.stride, D.threads, block,
            chunks, ranges.ptr, nb, window, None)
    for _ in range(3):
        H.call("honei_hip_product_smell_dv_staged", *args)
    H.synchronize()
    e0, e1 = H.Event(), H.Event()
    e0.record()
    for _ in range(reps):
        H.call("honei_hip_product_smell_dv_staged", *args)
    e1.record()
    ms = e0.elapsed_ms(e1) / reps
    alg = D.nnz * (vs + 4) + D.rows * 2 * vs
    gbs = alg / (ms * 1e-3) / 1e9
    log(f"[spmv] {label} (x staged in LDS, window {window}): {ms * 1e3:.1f} us/launch, {gbs:.0f} GB/s = "
        f"{100 * gbs / PEAK_GBS:.1f}%")
    return {"us_per_launch": round(ms * 1e3, 2), "algorithmic_bytes": int(alg), "achieved_gbs": round(gbs, 1),
            "pct_hbm_peak": round(100 * gbs / PEAK_GBS, 2), "window": int(window), "block": block, "chunks": chunks,
            "what": "honei_hip_product_smell_dv_staged, column windows per block computed once per matrix"}


def cg_q1_129(with_cpu):
    """CGSolver<..., NONE> and ConjugateGradients<..., JAC> (PCG-Jacobi) in
    FP64 on the Q1 129^2 operator (16,641 rows, ELL width 9), b = A x_true
    (x_true ~ U[-1,1), seed 1), x0 = 0, to relative residual 1e-8: operator,
    dinv and b built on the device; with_cpu times the oracle's C
    restatement of both solvers on one host core beside them."""
    side = 129
    D = device_q1(side, H.F64, with_dinv=True)
    n = D.rows
    xt = uniform_pm1(1, n)
    dxt = H.DeviceArray.from_host(xt)
    db = H.DeviceArray(n, np.float64)
    H.call("honei_hip_product_smell_dv_double", dxt.ptr, db.ptr, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, 0, n, D.ncpr,
           D.stride, 0, 1, None)
    dx = H.DeviceArray(n, np.float64)
    out = {}
    for name, solve in (("cg", lambda: H.cg_solve(D, db, dx, max_iters=10000, eps=1e-8)),
                        ("pcg_jacobi", lambda: H.pcg_jacobi(D, db, dx, D.dinv, max_iters=10000, eps=1e-8))):
        ts, its = [], 0
        for k in range(6):
            dx.zero()
            H.synchronize()
            t0 = time.perf_counter()
            its, _fd = solve()
            if k:
                ts.append(time.perf_counter() - t0)
        err = float(np.abs(dx.get() - xt).max())
        out[name] = {"gpu_ms": round(1e3 * float(np.median(ts)), 3), "iterations": int(its),
                     "max_err_vs_x_true": err}
    out["what"] = ("Q1 129^2 FP64 to 1e-8; GPU: honei_hip_cg_solve (CGSolver) and "
                   "honei_hip_pcg_jacobi (ConjugateGradients<JAC>, dinv = 1/diag)")
    if with_cpu:
        import oracle
        A = oracle.q1_matrix(side, np.float64)
        b = oracle.product(A, xt)
        dinv = 1.0 / oracle.q1_bands(side, np.float64)[4]
        for name, solve in (("cg", lambda: oracle.cg_solver(A, b, eps=1e-8)),
                            ("pcg_jacobi", lambda: oracle.pcg_jacobi(A, b, dinv, eps=1e-8))):
            cs = []
            for _ in range(3):
                t0 = time.perf_counter()
                _x, cits = solve()
                cs.append(time.perf_counter() - t0)
            out[name]["cpu_port_ms"] = round(1e3 * float(np.median(cs)), 3)
            out[name]["cpu_iterations"] = int(cits)
    for name in ("cg", "pcg_jacobi"):
        log(f"[{name} 129^2] GPU {out[name]['gpu_ms']} ms ({out[name]['iterations']} iterations), "
            f"CPU port {out[name].get('cpu_port_ms')} ms")
    return out


def pcg_iteration_rate(side=2049, iters=200):
    """PCG-Jacobi iterations in FP64 on the Q1 side^2 operator (device
    assembly, b = A 1): `iters` iterations with an unreachable tolerance,
    time per iteration and its algorithmic bytes 12 nnz + 16 n (t0 = A d, d
    gathered once) + 56 n (x, r update with z.r, r.r) + 32 n (direction)."""
    D = device_q1(side, H.F64, with_dinv=True)
    n = D.rows
    ones = H.DeviceArray.from_host(np.ones(n))
    db = H.DeviceArray(n, np.float64)
    H.call("honei_hip_product_smell_dv_double", ones.ptr, db.ptr, D.Aj.ptr, D.Ax.ptr, D.Arl.ptr, 0, n, D.ncpr,
           D.stride, 0, 1, None)
    dx = H.DeviceArray(n, np.float64)
    ts = []
    for k in range(4):
        dx.zero()
        H.synchronize()
        t0 = time.perf_counter()
        used, _ = H.pcg_jacobi(D, db, dx, D.dinv, max_iters=iters, eps=1e-30, check=False)
        if k:
            ts.append(time.perf_counter() - t0)
    per_it = float(np.median(ts)) / used
    nbytes = 12 * D.nnz + 104 * n
    out = {"side": side, "iterations": int(used), "us_per_iteration": round(1e6 * per_it, 2),
           "algorithmic_bytes_per_iteration": int(nbytes),
           "achieved_GBps": round(nbytes / per_it / 1e9, 1),
           "frac_of_8TBps": round(nbytes / per_it / 8e12, 3),
           "what": "PCG-Jacobi FP64 iteration (5 kernels: A d + d.t0, step, update + 2 dots, factor, "
                   "direction); whole solve call incl. defect, start and flag reads, divided by iterations"}
    log(f"[pcg {side}^2] {out['us_per_iteration']} us/iteration, {out['achieved_GBps']} GB/s")
    return out


def time_banded_q1(side, dtype, reps=50):
    """Banded Q1 product (BandedMatrixQx, no index stream): average launch
    time and algorithmic bytes 9n*s_v + n*(s_x + s_y).  The FP32 working set
    at 2049^2 (185 MB) would fit the 256 MB MALL, so launches alternate
    between two copies of the operands and every launch streams from HBM."""
    n = side * side
    sv = np.dtype(dtype).itemsize
    sfx = "float" if dtype == np.float32 else "double"
    sets = []
    for c in range(2):
        bands = [H.DeviceArray.from_host(uniform_pm1(10 + 9 * c + k, n).astype(dtype)) for k in range(9)]
        x = H.DeviceArray.from_host(uniform_pm1(2 + c, n).astype(dtype))
        y = H.DeviceArray(n, dtype)
        sets.append((bands, x, y, [b.ptr for b in bands] + [x.ptr, y.ptr, n, 0, side, None]))
    fn = f"honei_hip_product_bmdv_q1_{sfx}"
    for k in range(4):
        H.call(fn, *sets[k % 2][3])
    H.synchronize()
    e0, e1 = H.Event(), H.Event()
    e0.record()
    for k in range(reps):
        H.call(fn, *sets[k % 2][3])
    e1.record()
    ms = e0.elapsed_ms(e1) / reps
    alg = 9 * n * sv + 2 * n * sv
    gbs = alg / (ms * 1e-3) / 1e9
    log(f"[spmv] banded q1 {side}^2 {sfx}: {ms * 1e3:.1f} us/launch, {alg / 1e6:.1f} MB, {gbs:.0f} GB/s")
    del sets
    return {"us_per_launch": round(ms * 1e3, 2), "algorithmic_bytes": int(alg), "achieved_gbs": round(gbs, 1),
            "pct_hbm_peak": round(100 * gbs / PEAK_GBS, 2), "operand_copies": 2}


def read_traffic():
    """HBM bytes per launch of the roofline kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled)."""
    p = os.path.join(ROOT, "profiles", "ell_spmv_q1_2049_f32_traffic.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def batched_workload(nsys=NSYS, seed=BATCH_SEED):
    """Levels and device-resident coefficient / right-hand-side arrays of the
    batched workload (system i: seeds poisson_system_seeds(seed, i))."""
    levels = [int(v) for v in batch_levels(nsys, seed)]
    kap, rhs = [], []
    for i, L in enumerate(levels):
        side = 2 ** L + 1
        cs, bs = poisson_system_seeds(seed, i)
        k = H.DeviceArray((side - 1) ** 2, np.float64)
        b = H.DeviceArray(side * side, np.float64)
        H.fill_lognormal(k, 1.0, cs)
        H.fill_uniform(b, -1.0, 1.0, bs)
        kap.append(k)
        rhs.append(b)
    H.synchronize()
    return levels, kap, rhs


# Outer iterations per graph launch of the measured step: enough for the whole
# solve of the workload (median 12, max 17), so a step is one persistent graph
# launch plus one poll (the library default, 4, suits serving, where new
# requests join at launch boundaries).
STEP_ITERS_PER_LAUNCH = 24


def batched_pool(levels, **cfg):
    cfg.setdefault("iters_per_launch", STEP_ITERS_PER_LAUNCH)
    # repeated batches: each step's launch sized from the previous step's
    # largest iteration count (the first step runs STEP_ITERS_PER_LAUNCH)
    cfg.setdefault("first_launch_from_history", 1)
    lv = sorted(set(levels), reverse=True)
    counts = [levels.count(v) for v in lv]
    B = H.Batch(lv, counts, H.batch_config(**cfg))
    nxt = {v: 0 for v in lv}
    slots = []
    for L in levels:
        slots.append(B.slot(lv.index(L), nxt[L]))
        nxt[L] += 1
    return B, slots


def batched_step(B, slots, kap, rhs):
    """Assemble every system into its slot from the resident data, then run
    graph launches until all have retired.  Returns (seconds, per-system
    completion seconds, launches, slot infos)."""
    launches0 = B.launches()
    t0 = time.perf_counter()
    B.load_many(slots, kap, rhs)
    B.activate(True)
    done = np.full(len(slots), -1.0)
    while True:
        B.step()
        info = B.poll()
        t = time.perf_counter() - t0
        st = np.array([info[g].state for g in slots])
        newly = (st == 0) & (done < 0)
        done[newly] = t
        if (done >= 0).all():
            break
    return time.perf_counter() - t0, done, B.launches() - launches0, info


def fp64_single_solves(levels, kap, rhs, limit_s=60.0):
    """Plain single-system FP64 GPU solve: the same algorithm (CG
    preconditioned by one V(2,2) cycle, same rediscretised hierarchy, same
    tolerance) entirely in FP64, one system per call, host-driven; hierarchy
    build excluded from the time."""
    total, cycles, n_done, failed = 0.0, [], 0, 0
    for L, k, b in zip(levels, kap, rhs):
        mg = H.Multigrid(H.F64, 1, L)
        mg.configure(H.mg_config(tolerance=1e-8, max_outer=200))
        mg.build_q1_coeff(k, 0.8, 1)
        x = H.DeviceArray(b.n, np.float64)
        x.zero()
        H.synchronize()
        t0 = time.perf_counter()
        it, rd = mg.solve_pcg(b, x, check=False)
        total += time.perf_counter() - t0
        cycles.append(it)
        failed += not (rd < 1e-8)
        n_done += 1
        del mg
        if total > limit_s:
            break
    return total, n_done, cycles, failed


def serving_requests(seconds, rate=SERVE_RATE, seed=SERVE_SEED):
    """The serving stream (tests/patterns.py: Poisson arrivals, side uniform
    over 2^k+1 for k in SERVE_LEVELS, per-request seeds)."""
    return patterns.serving_requests(seconds, rate, seed, SERVE_LEVELS)


def q1_nnz(side):
    """Stored entries of the Q1 operator: 9 per interior row, 1 per Dirichlet row."""
    return 9 * (side - 2) ** 2 + 4 * (side - 1)


def slot_solve_bytes(L, iters, vs=4, fused=4, pair_from=5, mf=True):
    """Algorithmic bytes of every launch one slot of finest level L takes
    part in during a batched step with `iters` outer (CG) iterations, each
    launch counted by its own compulsory traffic (SURVEY-style: values and
    vectors once per launch, index streams only where a launch reads one).
    mf (the pool's default, matrix_free 1): the paired passes and the FP64
    row sums read the cell coefficients (8 bytes per cell) instead of
    stored values, and those values are never built.

      build    : coefficient read from its source and written, every coarser
                 coefficient from four finer cells, every stored level's
                 values (vs bytes each) and every level's dinv, the FP64
                 operator (not with mf), b copied, x0;
      settle   : r = b - A x in FP64 at the start and at the verification
                 (nnz*8 + 24n each, mf: 8 per cell + 24n; neighbours from
                 the grid, no indices), each followed by r/|r| -> FP32
                 (12n);
      per iteration, on the finest level n = n_L:
                 r.z (12n), p = z + beta p with q = A p and p.q in one
                 launch (nnz*8 + 28n, mf: 8 per cell + 28n: z and p read,
                 the new p and q written), x, r update with |r|^2 and the
                 next cycle's FP32 right-hand side (52n);
      per iteration, per level l (n_l, nnz_l):
                 l >= pair_from: paired Jacobi0 + defect and paired two
                 sweeps (nnz_l*vs + 16 n_l each, mf: 8 per cell + 16 n_l),
                 restriction (4 n_l + 4 n_{l-1}), the prolongation fused
                 into the paired sweeps (4 n_{l-1} more: the coarse result
                 read, the fine x read in place of the corrected vector);
                 2 <= l <= fused: the fused coarse launch's passes, each
                 with its index stream: Jacobi0, one Jacobi sweep, the
                 defect, two post sweeps (nnz_l*(vs+4) + 16 n_l, the defect
                 12 n_l), restriction and prolongation as above;
                 level 1 (3x3, coarse CG): not counted."""
    def n_(l):
        return (2 ** l + 1) ** 2

    def nnz_(l):
        return q1_nnz(2 ** l + 1)

    def cells(l):
        return 4 ** l

    n, nnz = n_(L), nnz_(L)
    op64 = 8 * cells(L) if mf else nnz * 8     # bytes of the FP64 operator a row sum reads
    build = 16 * cells(L) + 24 * n + (0 if mf else nnz * 8)
    for l in range(1, L + 1):
        stored = not (mf and l >= pair_from)
        build += 8 * cells(l) + (nnz_(l) * vs if stored else 0) + 4 * n_(l)
        if l < L:
            build += 8 * cells(l + 1) + 8 * cells(l)
    settle = 2 * (op64 + 24 * n + 12 * n)
    it = 12 * n + (op64 + 28 * n) + 52 * n
    for l in range(2, L + 1):
        nl, zl, nc = n_(l), nnz_(l), n_(l - 1)
        transfers = (4 * nl + 4 * nc) + (8 * nl + 4 * nc)
        if l >= pair_from:
            it += 2 * ((8 * cells(l) if mf else zl * vs) + 16 * nl) + (4 * nl + 4 * nc) + 4 * nc
        elif l <= fused:
            sweep = zl * (vs + 4) + 16 * nl
            it += 4 * sweep + (zl * (vs + 4) + 12 * nl) + transfers
    return build + settle + iters * it


def step_roofline(levels, iters, seconds, vs=4):
    """Step-level roofline: the algorithmic bytes of every launch of one
    batched step (slot_solve_bytes summed over the systems with their own
    iteration counts) over the step's wall time, against 8 TB/s."""
    total = sum(slot_solve_bytes(L, int(k), vs) for L, k in zip(levels, iters))
    gbs = total / seconds / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / PEAK_GBS, 4), "bytes_per_step": int(total),
            "seconds_per_step": round(seconds, 6),
            "what": ("algorithmic bytes of every launch of the step (hierarchy build, FP64 outer CG ops, every "
                     "V-cycle level's passes and transfers; per system from its outer iteration count) over "
                     "the step's wall time, against 8 TB/s")}


def device_fields(L, cs, bs):
    """Host copies of the coefficient and right-hand side the device draws
    from a system's seeds (the fields the batch and the serving loop solve)."""
    side = 2 ** L + 1
    k = H.DeviceArray((side - 1) ** 2, np.float64)
    b = H.DeviceArray(side * side, np.float64)
    H.fill_lognormal(k, 1.0, cs)
    H.fill_uniform(b, -1.0, 1.0, bs)
    H.synchronize()
    return k.get(), b.get()


def sample_batch_results(B, levels, slots, info, per_level=ERR_SAMPLE):
    """(level, coefficient, b, x, device relative residual) of the first
    `per_level` systems of each size in workload order, for the FP64
    direct-solve error of the CPU leg."""
    out, seen = [], {}
    for i, (L, g) in enumerate(zip(levels, slots)):
        if seen.get(L, 0) >= per_level:
            continue
        seen[L] = seen.get(L, 0) + 1
        x = H.DeviceArray((2 ** L + 1) ** 2, np.float64)
        B.result(g, x)
        kappa, b = device_fields(L, *poisson_system_seeds(BATCH_SEED, i))
        out.append((L, kappa, b, x.get(), float(info[g].rel)))
    return out


def level_launch_rates(B, levels, level, reps=20):
    """The step's own launch shape at full activity: honei_hip_batch_time_level
    over every slot present on `level` (all systems of that size and larger),
    and the FP64 q = A p over every slot; bytes by the rules of
    slot_solve_bytes.  Overwrites the pool's vectors (after the results are
    taken)."""
    pre, rst, pro, post, spmv = B.time_level(level, reps)
    side = 2 ** level + 1
    n, nc = side * side, (2 ** (level - 1) + 1) ** 2
    cells = (side - 1) ** 2
    cap = sum(1 for L in levels if L >= level)
    # the pool's default weight source (matrix_free 1): 8 bytes per cell
    # instead of the stored values
    rows = {"paired_pre": (pre, cap * (8 * cells + 16 * n),
                           "k_b_pair<float,PRE,weights from cells>: Jacobi0 then the defect"),
            "restriction": (rst, cap * (4 * n + 4 * nc), "k_b_restrict_rows"),
            "prolongation": (pro, cap * (8 * n + 4 * nc), "k_b_prolong_rows: t = x + P e"),
            # levels that pair fuse the prolongation into the post-smoothing
            # launch (pro = 0 there): the coarse result is read once more
            "paired_post": (post, cap * (8 * cells + 16 * n + (4 * nc if pro == 0 else 0)),
                            "k_b_pair<float,POST,weights from cells%s>: two Jacobi sweeps"
                            % (", prolongation fused" if pro == 0 else ""))}
    out = {"level": level, "side": side, "slots": cap}
    for key, (us, nb, what) in rows.items():
        if us > 0:
            g = nb / (us * 1e-6) / 1e9
            out[key] = {"kernel": what, "us": round(us, 2), "algorithmic_bytes": int(nb), "achieved_gbs": round(g, 1),
                        "frac": round(g / PEAK_GBS, 4)}
    if spmv > 0:
        # z (4) and p (8) read, the new p and q (8 each) written, 8 per cell
        nb = sum(8 * 4 ** L + 28 * (2 ** L + 1) ** 2 for L in levels)
        g = nb / (spmv * 1e-6) / 1e9
        out["outer_spmv_dot"] = {"kernel": ("k_b_dir_spmv<weights from cells, float>: FP64 p = z + beta p, q = A p, "
                                            "p.q, every slot"),
                                 "us": round(spmv, 2),
                                 "algorithmic_bytes": int(nb), "achieved_gbs": round(g, 1),
                                 "frac": round(g / PEAK_GBS, 4)}
    out["what"] = ("launches of the batched step's shape over every slot of the level (full activity), timed "
                   "back to back with HIP events on the pool's stream after the timed steps")
    log("[batch] level %d launches over %d slots: " % (level, cap)
        + ", ".join(f"{k} {v['us']} us ({v['frac']:.2f})" for k, v in out.items() if isinstance(v, dict)))
    return out


def serving_payload(seconds, sources=8):
    """The serving stream with a host payload: each request's coefficient and
    right-hand side are copied from pinned host memory when it is admitted
    (`sources` distinct systems per size, request i of a size taking source
    i mod `sources`), and its solution is copied back to pinned host memory
    (one buffer per size) before the request counts as done; latency covers
    both copies."""
    reqs = serving_requests(seconds)
    src, xbuf, seen = {}, {}, {}
    for L in SERVE_LEVELS:
        src[L] = []
        for k in range(sources):
            cs, bs = poisson_system_seeds(SERVE_SEED + 100, 16 * L + k)
            kappa, b = device_fields(L, cs, bs)
            pk, pb = H.PinnedArray(kappa.size), H.PinnedArray(b.size)
            pk.a[:] = kappa
            pb.a[:] = b
            src[L].append((pk, pb))
        xbuf[L] = H.PinnedArray((2 ** L + 1) ** 2)
    mix = []
    for q in reqs:
        k = seen.get(q[1], 0)
        seen[q[1]] = k + 1
        pk, pb = src[q[1]][k % sources]
        mix.append(q + (None, pk, pb, xbuf[q[1]]))
    _res, st = H.serve(SERVE_LEVELS, mix, budget_bytes=SERVE_BUDGET)
    mb = sum(24 * (2 ** q[1] + 1) ** 2 for q in reqs) / 1e6
    out = {"requests": len(reqs), "completed": st.completed, "not_converged": st.not_converged,
           "solves_per_s": round(st.solves_per_s, 2), "wall_s": round(st.wall_s, 3),
           "latency_ms": {"p50": round(st.p50_ms, 2), "p99": round(st.p99_ms, 2), "max": round(st.max_ms, 2)},
           "host_link_MB": round(mb, 1), "host_link_GBps": round(mb / 1e3 / max(st.wall_s, 1e-9), 2),
           "what": (f"the serving stream for {seconds:.0f} s with coefficient and right-hand side copied from pinned "
                    "host memory at admission and the solution copied back to pinned host memory; latency from "
                    "arrival to the end of the return copy")}
    log(f"[serve payload] {st.completed} solves in {st.wall_s:.2f} s ({st.solves_per_s:.1f}/s), p50 {st.p50_ms:.1f} ms "
        f"p99 {st.p99_ms:.1f} ms, {out['host_link_GBps']} GB/s over the host link")
    del src, xbuf
    return out


def serving_capacity(seconds, rates=(2000.0, 3000.0, 4000.0, 6000.0, 9000.0)):
    """Open-loop serving streams of increasing arrival rate (same size mix,
    seeds and 64 GB budget as the serving run, `seconds` long each):
    throughput and p50/p99 per rate; the knee is the first rate whose
    throughput falls below 95 % of the offered rate, the capacity the
    largest throughput measured."""
    rows = []
    for rate in rates:
        reqs = serving_requests(seconds, rate=rate, seed=SERVE_SEED)
        _res, st = H.serve(SERVE_LEVELS, reqs, budget_bytes=SERVE_BUDGET)
        rows.append({"arrival_rate": rate, "requests": len(reqs), "solves_per_s": round(st.solves_per_s, 1),
                     "p50_ms": round(st.p50_ms, 2), "p99_ms": round(st.p99_ms, 2), "max_queue": st.max_queue,
                     "not_converged": st.not_converged})
        log(f"[capacity] {rate:.0f} req/s: {st.solves_per_s:.0f} solves/s, p50 {st.p50_ms:.1f} ms, "
            f"p99 {st.p99_ms:.1f} ms, max queue {st.max_queue}")
        if st.solves_per_s < 0.8 * rate:
            break
    knee = next((r["arrival_rate"] for r in rows if r["solves_per_s"] < 0.95 * r["arrival_rate"]), None)
    return {"streams": rows, "knee_arrival_rate": knee,
            "capacity_solves_per_s": max(r["solves_per_s"] for r in rows), "seconds_per_stream": seconds,
            "what": "open-loop Poisson streams at increasing rates; wall time includes draining the queue"}


# ---- CPU leg: the reference's own CPU path and the FP64 oracle --------------------------
def cpu_workers():
    """Host processes of the all-cores CPU leg: the CPUs this process may run
    on, at most 16 (the host share of a one-GPU job on the MI355X boxes)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, aff))


def _reference_cpu_job(q):
    try:
        q.put(reference_cpu_path())
    except Exception as e:   # reported in the JSON line, never fatal to the bench
        q.put({"error": repr(e)})


def start_reference_cpu_path():
    """reference_cpu_path in a forked host process (started before the GPU is
    touched; it runs on one core beside the GPU work)."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    q = ctx.Queue()
    p = ctx.Process(target=_reference_cpu_job, args=(q,))
    p.start()
    return p, q


def host_cpu_info():
    """Host CPU model, logical CPUs and the CPUs this process may run on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"model": model, "nproc": os.cpu_count(), "affinity": aff}


def reference_cpu_path():
    """The reference's own CPU solvers (oracle/_ref: HONEI's tags::CPU code
    built from its sources), one host core: CGSolver<CPU, NONE> on the Q1
    129^2 FP64 system (b = A x_true, seed 1, x0 = 0, to 1e-8) and
    Multigrid<CPU, CPU, NONE, JAC, V, MIXED> on Q1 2049^2 (b = A x_true,
    seed 3, 11 grids, V(2,2), one V-cycle per FP64 outer step, to 1e-10),
    median wall time of the solve call.  None where oracle/_ref is absent."""
    from oracle import ref
    if not ref.available():
        return None
    import oracle
    out = {"kind": "reference", "cores": 1}
    A = ref.q1_ell(129)
    xt = uniform_pm1(1, 129 * 129)
    b = oracle.product(oracle.q1_matrix(129), xt)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        x, used = ref.cg_solver(A, b, np.zeros_like(b), 10000, 1e-8)
        ts.append(time.perf_counter() - t0)
    out["cg_q1_129_fp64"] = {"ms": round(1e3 * float(np.median(ts)), 3), "iterations": int(used),
                             "max_err_vs_x_true": float(np.abs(x - xt).max()),
                             "what": "CGSolver<tags::CPU, NONE> of the reference build"}
    log(f"[cpu ref] CG 129^2: {out['cg_q1_129_fp64']['ms']} ms, {used} iterations")
    xt = uniform_pm1(3, SIDE * SIDE)
    b = oracle.product(oracle.q1_matrix(SIDE), xt)
    ts, cyc = [], 0
    for _ in range(2):
        x, cyc, _inner, secs = ref.mg_mixed(LEVEL, b, np.zeros_like(b), n_max_iter=1, n_pre=2, tol=1e-10, timed=True)
        ts.append(secs)
    out["mixed_mg_q1_2049"] = {"ms": round(1e3 * float(np.median(ts)), 1), "outer_iterations": int(cyc),
                               "max_err_vs_x_true": float(np.abs(x - xt).max()),
                               "what": ("Multigrid<tags::CPU, tags::CPU, NONE, JAC, V, MIXED> of the reference build, "
                                        "one V(2,2) cycle per outer step, to 1e-10; set-up not timed")}
    log(f"[cpu ref] MIXED MG 2049^2: {out['mixed_mg_q1_2049']['ms']} ms, {cyc} outer iterations")
    return out


def fp64_oracle_errors(samples, workers=16):
    """Final error of GPU results against the oracle's FP64 direct solve
    (sparse LU plus refinement to 1e-14, oracle.direct_solve) of the same
    systems, in fresh host processes (no GPU state): per group, the largest
    max|x - x64| / max|x64| and the oracle's own FP64 relative residual of
    the GPU solution."""
    import multiprocessing as mp
    jobs = [(g, L, k, b, x, rel) for g, rows in samples.items() for (L, k, b, x, rel) in rows]
    if not jobs:
        return {}
    with mp.get_context("spawn").Pool(max(1, min(workers, len(jobs)))) as pool:
        res = pool.map(_direct_error_job, jobs)
    out = {}
    for (g, L, _k, _b, _x, rel), (err, orel) in zip(jobs, res):
        o = out.setdefault(g, {"max_err_vs_fp64_oracle": 0.0, "max_oracle_rel_residual": 0.0, "per_side": {}})
        o["max_err_vs_fp64_oracle"] = max(o["max_err_vs_fp64_oracle"], err)
        o["max_oracle_rel_residual"] = max(o["max_oracle_rel_residual"], orel)
        side = str(2 ** L + 1)
        o["per_side"][side] = max(o["per_side"].get(side, 0.0), err)
    for g, o in out.items():
        o["systems"] = sum(1 for j in jobs if j[0] == g)
        log(f"[cpu fp64] {g}: max error vs FP64 direct solve {o['max_err_vs_fp64_oracle']:.2e} over "
            f"{o['systems']} systems, oracle residual <= {o['max_oracle_rel_residual']:.2e}")
    return out


def _direct_error_job(job):
    import oracle
    _g, L, kappa, b, x, _rel = job
    side = 2 ** L + 1
    A = oracle.q1_coeff_matrix(side, kappa)
    x64, _r, _s = oracle.direct_solve(A, b)
    guess = np.zeros_like(b)
    bd = np.zeros(side * side, bool)
    bd[:side] = bd[-side:] = True
    bd[::side] = True
    guess[bd] = b[bd]
    r0 = np.linalg.norm(oracle.defect(A, b, guess))
    orel = float(np.linalg.norm(oracle.defect(A, b, x)) / r0)
    return float(np.abs(x - x64).max() / np.abs(x64).max()), orel


def batch_kernel_roofline(B, level):
    """Bandwidth of the batched Jacobi sweep on `level` from the autotuner's
    own timing (tuning batch of `slots` systems sharing one index array)."""
    T, block, group, us, ns = B.tuning(level)
    if us <= 0:
        return None
    side = 2 ** level + 1
    n = side * side
    nnz = 9 * (side - 2) ** 2 + 4 * (side - 1)
    # per system: FP32 values + x, rhs, dinv, out per row; once: the shared indices and row lengths
    bytes_ = ns * (nnz * 4 + n * 4 * 4) + nnz * 4 + n * 4
    gbs = bytes_ / (us * 1e-6) / 1e9
    out = {"row_sweep": {"kernel": f"k_b_ell<float,{T},JACOBI,{group}> level {level} ({side}^2), block {block}",
                         "slots": ns, "us": round(us, 2), "algorithmic_bytes": int(bytes_),
                         "achieved_gbs": round(gbs, 1), "frac": round(gbs / PEAK_GBS, 4),
                         "bytes_rule": "slots*(nnz*4 + n*16) + nnz*4 + n*4: the column indices are shared by all "
                                       "slots",
                         "note": "one pass per launch; levels from 33^2 up run the paired launches below"}}
    pre, post = B.pair_timing(level)   # the pool's weight source: the cell coefficients
    pbytes = ns * (8 * (side - 1) ** 2 + n * 16)
    for name, t, what in (("paired_pre", pre, "k_b_pair<float,PRE,weights from cells>: Jacobi0 then the defect"),
                          ("paired_post", post, "k_b_pair<float,POST,weights from cells>: two Jacobi sweeps")):
        if t > 0:
            g = pbytes / (t * 1e-6) / 1e9
            out[name] = {"kernel": f"{what}, level {level} ({side}^2)", "slots": ns, "us": round(t, 2),
                         "algorithmic_bytes": int(pbytes), "achieved_gbs": round(g, 1),
                         "frac": round(g / PEAK_GBS, 4),
                         "bytes_rule": "slots*(8*cells + n*16): the cell coefficients once for both passes, rhs, "
                                       "dinv, the input or first output, the output; no index or value stream"}
    return out


def cpu_baseline_batched(levels, budget_s=15.0):
    """The clean-room C restatement of the CPU MIXED multigrid (one host
    core) on the batched workload's own systems, taken in workload order
    (same seeds, coefficient and right-hand side drawn on the host) until
    `budget_s` seconds of solve time are spent or all are solved.  Only the
    solve is timed; building each host hierarchy is not."""
    import oracle
    from tests.patterns import lognormal_cells
    spent, done, conv, outer_sum = 0.0, 0, 0, 0
    per = {}
    for i, L in enumerate(levels):
        side = 2 ** L + 1
        cs, bs = poisson_system_seeds(BATCH_SEED, i)
        kappa = lognormal_cells(cs, (side - 1) ** 2)
        b = uniform_pm1(bs, side * side)
        Hh = oracle.q1_coeff_hierarchy(L, kappa, averaging=1, omega=0.8)
        A64 = oracle.q1_coeff_matrix(side, kappa)
        cfg = oracle.mg_config(1, L, n_max_iter=1, tolerance=1e-8, initial_zero=True, zero_inner_start=True,
                               max_outer=200)
        t0 = time.perf_counter()
        _, outer, _, hist = oracle.mg_mixed(Hh, cfg, A64, b)
        dt = time.perf_counter() - t0
        spent += dt
        done += 1
        outer_sum += outer
        conv += int(len(hist) > 0 and hist[-1] < 1e-8)
        t, c = per.get(L, (0.0, 0))
        per[L] = (t + dt, c + 1)
        if spent > budget_s:
            break
    log(f"[cpu] {done} systems in {spent:.2f} s of solve time, {conv} converged, "
        f"{outer_sum / max(done, 1):.1f} outer iterations on average")
    return {"value": round(done / spent, 4), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": (f"MIXED multigrid (FP32 V(2,2) inner, FP64 outer defect correction, tol 1e-8, at most 200 "
                       f"outer) of the oracle's C restatement on one host core: the first {done} of the "
                       f"{len(levels)} batched-workload systems in workload order ("
                       + ", ".join(f"{per[L][1]}x{2 ** L + 1}^2 {per[L][0]:.2f}s" for L in sorted(per))
                       + f"), {spent:.1f} s of solve time, {conv} reached the tolerance; "
                       f"host hierarchy build not timed")}


def _cpu_system(i, L):
    """Host hierarchy, FP64 operator, b and MIXED config of workload system i."""
    import oracle
    from tests.patterns import lognormal_cells
    side = 2 ** L + 1
    cs, bs = poisson_system_seeds(BATCH_SEED, i)
    kappa = lognormal_cells(cs, (side - 1) ** 2)
    b = uniform_pm1(bs, side * side)
    Hh = oracle.q1_coeff_hierarchy(L, kappa, averaging=1, omega=0.8)
    A64 = oracle.q1_coeff_matrix(side, kappa)
    cfg = oracle.mg_config(1, L, n_max_iter=1, tolerance=1e-8, initial_zero=True, zero_inner_start=True,
                           max_outer=200)
    return Hh, cfg, A64, b


def _cpu_worker(jobs, levels, barrier, q):
    import oracle
    systems = [_cpu_system(i, levels[i]) for i in jobs]
    barrier.wait()
    conv = 0
    for Hh, cfg, A64, b in systems:
        _, _, _, hist = oracle.mg_mixed(Hh, cfg, A64, b)
        conv += int(len(hist) > 0 and hist[-1] < 1e-8)
    q.put((len(systems), conv, time.time()))


def cpu_baseline_all_cores(levels, workers):
    """The same restatement on `workers` host processes at once (one system
    per process at a time): every system of the workload, assigned to the
    processes longest-first by unknowns; each process builds its host
    hierarchies, all start together at a barrier, and the wall time from
    the barrier to the last finish is timed.  Run before the GPU is touched
    (the processes are forked)."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    order = sorted(range(len(levels)), key=lambda i: -(2 ** levels[i] + 1) ** 2)
    jobs, load = [[] for _ in range(workers)], [0] * workers
    for i in order:
        w = load.index(min(load))
        jobs[w].append(i)
        load[w] += (2 ** levels[i] + 1) ** 2
    barrier, q = ctx.Barrier(workers + 1), ctx.Queue()
    procs = [ctx.Process(target=_cpu_worker, args=(jobs[w], levels, barrier, q)) for w in range(workers)]
    for p in procs:
        p.start()
    barrier.wait()
    t0 = time.time()
    res = [q.get() for _ in range(workers)]
    for p in procs:
        p.join()
    wall = max(r[2] for r in res) - t0
    done, conv = sum(r[0] for r in res), sum(r[1] for r in res)
    log(f"[cpu] {workers} processes: {done} systems in {wall:.2f} s wall, {conv} converged")
    return {"value": round(done / wall, 4), "unit": "solves/s", "cores": workers, "kind": "port",
            "sample": (f"MIXED multigrid (FP32 V(2,2) inner, FP64 outer defect correction, tol 1e-8, at most 200 "
                       f"outer) of the oracle's C restatement, all {done} batched-workload systems solved by "
                       f"{workers} host processes at once (longest first by unknowns), {wall:.1f} s wall from a "
                       f"common start; {conv} reached the tolerance; host hierarchy build not timed")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed batched solves of the 256 systems")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the Poisson arrival stream of the serving run (0: skip)")
    ap.add_argument("--capacity-seconds", type=float, default=6.0,
                    help="length of each stream of the serving capacity sweep (0: skip)")
    ap.add_argument("--payload-seconds", type=float, default=10.0,
                    help="length of the serving stream with host payloads (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the single-system FP64 comparison")
    ap.add_argument("--batch-only", action="store_true", help="only the batched step (for profiling)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="with --batch-only: enqueue the step's kernels directly instead of graph launches "
                         "(same kernels and results; for kernel-trace profiling)")
    ap.add_argument("--no-pair", action="store_true",
                    help="with --batch-only: one launch per smoothing pass (pair_sweeps = 0), for comparison")
    ap.add_argument("--weights", type=int, default=-1, choices=(-1, 0, 1, 2),
                    help="with --batch-only: weight source of the paired launches and FP64 outer row sums "
                         "(matrix_free 0 stored values, 1 cell coefficients, 2 distinct-weight arrays; "
                         "-1: the pool's default), for comparison")
    ap.add_argument("--iters-per-launch", type=int, default=0,
                    help="with --batch-only: outer iterations per graph launch (0: the pool's default)")
    ap.add_argument("--spmv-only", action="store_true", help="only the SpMV section")
    ap.add_argument("--pcg-only", action="store_true",
                    help="only the PCG-Jacobi iteration rate on Q1 2049^2 FP64 (for profiling)")
    ap.add_argument("--serve-only", action="store_true", help="only the serving stream (for profiling)")
    ap.add_argument("--pmc-spmv", action="store_true",
                    help="5 launches each of the Q1 2049^2 FP64 ELL SpMV, the random 4M ELL SpMV (FP32, FP64) "
                         "and its sliced copy, then 5 of k_scale<float> on 2^26 floats (for rocprofv3 --pmc)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="counter runs: 5 launches of the roofline SpMV and 5 of a 2^26-float scale kernel "
                         "(reads and writes 256 MiB each) that calibrates FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the path is single-GPU by design",
              file=sys.stderr)
        sys.exit(2)

    # the CPU legs fork host processes: start them before the GPU is touched
    cpu_all, ref_proc = None, None
    full = not (args.batch_only or args.serve_only or args.spmv_only or args.pmc_probe or args.pcg_only
                or args.pmc_spmv)
    if full and not args.no_cpu:
        os.environ["OMP_NUM_THREADS"] = "1"
        cpu_all = cpu_baseline_all_cores([int(v) for v in batch_levels(NSYS, BATCH_SEED)], cpu_workers())
        cpu_all["host"] = host_cpu_info()
        ref_proc = start_reference_cpu_path()   # one core, runs beside the GPU work
    H.init(0)
    n = SIDE * SIDE
    if args.pmc_probe:
        D32 = device_q1(SIDE, H.F32)
        x32 = H.DeviceArray.from_host(uniform_pm1(2, n).astype(np.float32))
        y32 = H.DeviceArray(n, np.float32)
        for _ in range(5):
            H.call("honei_hip_product_smell_dv_float", x32.ptr, y32.ptr, D32.Aj.ptr, D32.Ax.ptr, D32.Arl.ptr, 0, n,
                   9, D32.stride, 0, 1, None)
        big = H.DeviceArray(1 << 26, np.float32)
        big.zero()
        for _ in range(5):
            H.call("honei_hip_scale_one", H.F32, big.ptr, 1.0, 1 << 26, None)
        H.synchronize()
        print(json.dumps({"pmc_probe": "done"}))
        return

    if args.pmc_spmv:
        D64 = device_q1(SIDE, H.F64)
        x64 = H.DeviceArray.from_host(uniform_pm1(2, n))
        y64 = H.DeviceArray(n, np.float64)
        for _ in range(5):
            H.call("honei_hip_product_smell_dv_double", x64.ptr, y64.ptr, D64.Aj.ptr, D64.Ax.ptr, D64.Arl.ptr, 0, n,
                   9, D64.stride, 0, 1, None)
        H.synchronize()
        del D64, x64, y64
        for vt, dt in ((H.F32, np.float32), (H.F64, np.float64)):
            R = device_random_ell(1 << 22, vt, seed=2)
            xr = H.DeviceArray.from_host(uniform_pm1(2, R.rows).astype(dt))
            yr = H.DeviceArray(R.rows, dt)
            for _ in range(5):
                H.call("honei_hip_product_smell_dv_float" if vt == H.F32 else "honei_hip_product_smell_dv_double",
                       xr.ptr, yr.ptr, R.Aj.ptr, R.Ax.ptr, R.Arl.ptr, 0, R.rows, R.ncpr, R.stride, 0, 1, None)
            S = H.Sell(R)
            for _ in range(5):
                S.product(yr, xr)
            H.synchronize()
            del S, R, xr, yr
        big = H.DeviceArray(1 << 26, np.float32)
        big.zero()
        for _ in range(5):
            H.call("honei_hip_scale_one", H.F32, big.ptr, 1.0, 1 << 26, None)
        H.synchronize()
        print(json.dumps({"pmc_spmv": "done"}))
        return

    if args.pcg_only:
        print(json.dumps({"pcg_only": pcg_iteration_rate()}))
        return

    if args.batch_only:
        levels, kap, rhs = batched_workload()
        kw = dict(use_graphs=0 if args.no_graphs else 1, pair_sweeps=0 if args.no_pair else 1)
        if args.weights >= 0:
            kw["matrix_free"] = args.weights
        if args.iters_per_launch > 0:
            kw["iters_per_launch"] = args.iters_per_launch
        B, slots = batched_pool(levels, **kw)
        for _ in range(args.warmup):
            batched_step(B, slots, kap, rhs)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sec, done, nl, info = batched_step(B, slots, kap, rhs)
        total = time.perf_counter() - t0
        print(json.dumps({"batch_only": {"solves_per_s": round(len(levels) * args.steps / total, 2),
                                         "steps": args.steps, "launches_last_step": nl,
                                         "shapes": {L: B.tuning(L) for L in range(2, max(levels) + 1)},
                                         "pair_us": {L: B.pair_timing(L) for L in range(2, max(levels) + 1)}}}))
        return

    if args.serve_only:
        reqs = serving_requests(args.serve_seconds)
        res, st = H.serve(SERVE_LEVELS, reqs, budget_bytes=SERVE_BUDGET)
        print(json.dumps({"serve_only": {"requests": len(reqs), "completed": st.completed,
                                         "solves_per_s": round(st.solves_per_s, 2), "p50_ms": round(st.p50_ms, 2),
                                         "p99_ms": round(st.p99_ms, 2), "launches": st.launches}}))
        return

    # ---- standalone SpMV (roofline kernel: Q1 2049^2, FP32, T=1) ----------
    spmv = {}
    for vt, dt, key in ((H.F32, np.float32, "q1_2049_f32"), (H.F64, np.float64, "q1_2049_f64")):
        sets = [(with_idx64(device_q1(SIDE, vt)), H.DeviceArray.from_host(uniform_pm1(2 + c, n).astype(dt)),
                 H.DeviceArray(n, dt)) for c in range(2)]
        spmv[key] = time_spmv(sets, 50, f"q1 2049^2 {key[-3:]}")
        spmv[key + "_idx64"] = time_spmv(sets, 50, f"q1 2049^2 {key[-3:]}, 64-bit indices", idx64=True)
        spmv[key + "_idx64"]["what"] = ("honei_hip_product_smell_dv_*_idx64 on the reference's unsigned long Aj/Arl; "
                                        "bytes with 8-byte indices")
        del sets
        if args.spmv_only and vt == H.F32:
            print(json.dumps({"spmv": spmv}))
            return
    spmv["q1_2049_banded_f32"] = time_banded_q1(SIDE, np.float32)
    spmv["q1_2049_banded_f64"] = time_banded_q1(SIDE, np.float64)
    for vt, dt, key in ((H.F32, np.float32, "random_4m_f32"), (H.F64, np.float64, "random_4m_f64")):
        sets = []
        for c in range(2):
            R = device_random_ell(1 << 22, vt, seed=2 + c)
            sets.append((R, H.DeviceArray.from_host(uniform_pm1(2 + c, R.rows).astype(dt)), H.DeviceArray(R.rows, dt)))
        spmv[key] = time_spmv(sets, 30, key)
        spmv[key]["nnz"] = sets[0][0].nnz
        spmv[key + "_sliced"] = time_sell(sets, 30, key)
        spmv[key + "_sliced_global_x"] = time_sell(sets, 30, key, stage_x=False)
        R, xr, yr = sets[0]
        spmv[key + "_staged"] = time_spmv_staged(R, 30, xr, yr, key)
        del sets, R, xr, yr
    spmv["hbm_ceiling"] = hbm_ceiling()
    cg129 = cg_q1_129(not args.no_cpu)
    pcg2049 = pcg_iteration_rate()
    if not args.no_cpu:
        spmv["cpu_q1_2049_f32"] = cpu_spmv_q1(SIDE)

    # ---- batched solve: the measured step ------------------------------------
    levels, kap, rhs = batched_workload()
    B, slots = batched_pool(levels)
    log(f"[batch] {len(levels)} systems, pool {B.device_bytes / 2**30:.2f} GiB, shapes "
        + ", ".join(f"L{L}:{B.shape(L)}" for L in sorted(set(levels))))
    for _ in range(args.warmup):
        batched_step(B, slots, kap, rhs)
    H.synchronize()
    step_s, lat, launches = [], [], 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        sec, done, nl, info = batched_step(B, slots, kap, rhs)
        step_s.append(sec)
        lat.extend(done.tolist())
        launches += nl
    H.synchronize()
    total = time.perf_counter() - t_start
    iters = np.array([info[g].iters for g in slots])
    rels = np.array([info[g].rel for g in slots])
    status = np.array([info[g].status for g in slots])
    log(f"[batch] {args.steps} steps in {total:.3f} s: {len(levels) * args.steps / total:.1f} solves/s; "
        f"outer iterations median {int(np.median(iters))} max {iters.max()}; max rel residual {rels.max():.2e}; "
        f"{launches / args.steps:.1f} graph launches per step")
    value = len(levels) * args.steps / total
    lat_ms = np.array(lat) * 1e3
    step_rl = step_roofline(levels, iters, total / args.steps)
    log(f"[batch] step roofline: {step_rl['bytes_per_step'] / 1e9:.1f} GB of launches per step, "
        f"{step_rl['achieved']:.0f} GB/s = {step_rl['frac']:.3f} of peak")
    err_samples = {"batch": sample_batch_results(B, levels, slots, info)}

    # ---- plain single-system FP64 GPU solve of the same systems -------------------
    fp64 = None
    if not args.no_fp64:
        t64, n64, cyc, f64fail = fp64_single_solves(levels, kap, rhs)
        fp64 = {"solves_per_s": round(n64 / t64, 3), "systems": n64, "seconds": round(t64, 3),
                "iterations_median": int(np.median(cyc)), "iterations_max": int(max(cyc)),
                "not_converged": f64fail,
                "what": ("honei_hip_mg_solve_pcg with an FP64 hierarchy: FP64 CG preconditioned by an FP64 "
                         "V(2,2) cycle, one system per call, host-driven, same tolerance")}
        log(f"[fp64] {n64} single-system FP64 solves in {t64:.3f} s ({n64 / t64:.1f} solves/s)")
    batch_kernel = batch_kernel_roofline(B, max(levels))
    batch_kernel["step_shape"] = level_launch_rates(B, levels, max(levels))
    del B
    # the same step with BF16-stored V-cycle matrices (FP32 arithmetic)
    Bb, slots_b = batched_pool(levels, inner_vtype=H.BF16)
    batched_step(Bb, slots_b, kap, rhs)
    tb = []
    for _ in range(max(1, args.steps)):
        sec, _done, _nl, info_b = batched_step(Bb, slots_b, kap, rhs)
        tb.append(sec)
    it_b = np.array([info_b[g].iters for g in slots_b])
    err_samples["bf16_inner"] = sample_batch_results(Bb, levels, slots_b, info_b)
    bf16 = {"solves_per_s": round(len(levels) / float(np.median(tb)), 2),
            "outer_iterations_median": int(np.median(it_b)), "outer_iterations_max": int(it_b.max()),
            "max_rel_residual": float(max(info_b[g].rel for g in slots_b)),
            "all_converged": bool(all(info_b[g].status == 0 for g in slots_b))}
    log(f"[batch bf16] {bf16['solves_per_s']} solves/s, outer median {bf16['outer_iterations_median']}")
    del Bb
    # the same batched step entirely in FP64 (FP64 V-cycle values and vectors):
    # separates the precision gain from the batching gain
    Bd, slots_d = batched_pool(levels, inner_vtype=H.F64)
    batched_step(Bd, slots_d, kap, rhs)
    td = []
    for _ in range(max(1, args.steps)):
        sec, _done, _nl, info_d = batched_step(Bd, slots_d, kap, rhs)
        td.append(sec)
    it_d = np.array([info_d[g].iters for g in slots_d])
    err_samples["fp64_batched"] = sample_batch_results(Bd, levels, slots_d, info_d)
    fp64b = {"solves_per_s": round(len(levels) / float(np.median(td)), 2),
             "outer_iterations_median": int(np.median(it_d)), "outer_iterations_max": int(it_d.max()),
             "max_rel_residual": float(max(info_d[g].rel for g in slots_d)),
             "all_converged": bool(all(info_d[g].status == 0 for g in slots_d)),
             "what": ("the same 256-system batched step (one pool, same launches, CG outer) with an FP64 "
                      "V-cycle: FP64 values and vectors on every level")}
    log(f"[batch fp64] {fp64b['solves_per_s']} solves/s, outer median {fp64b['outer_iterations_median']}")
    del Bd

    # ---- serving stream ---------------------------------------------------------------
    serving = None
    if args.serve_seconds > 0:
        reqs = serving_requests(args.serve_seconds)
        log(f"[serve] {len(reqs)} requests over {args.serve_seconds:.0f} s, levels {SERVE_LEVELS}, "
            f"budget {SERVE_BUDGET >> 30} GiB")
        # the first ERR_SAMPLE requests of each size return their solution
        picked, seen = [], {}
        for i, q in enumerate(reqs):
            if seen.get(q[1], 0) < ERR_SAMPLE:
                seen[q[1]] = seen.get(q[1], 0) + 1
                picked.append(i)
        xs = {i: H.DeviceArray((2 ** reqs[i][1] + 1) ** 2, np.float64) for i in picked}
        res, st = H.serve(SERVE_LEVELS, [q + (xs[i],) if i in xs else q for i, q in enumerate(reqs)],
                          budget_bytes=SERVE_BUDGET)
        err_samples["serving"] = [(reqs[i][1], *device_fields(reqs[i][1], reqs[i][2], reqs[i][3]), xs[i].get(),
                                   float(res[i].rel)) for i in picked]
        del xs
        serving = {"requests": len(reqs), "completed": st.completed, "not_converged": st.not_converged,
                   "rejected": st.rejected, "wall_s": round(st.wall_s, 3),
                   "solves_per_s": round(st.solves_per_s, 2),
                   "latency_ms": {"p50": round(st.p50_ms, 2), "p99": round(st.p99_ms, 2),
                                  "max": round(st.max_ms, 2), "mean": round(st.mean_ms, 2)},
                   "max_queue": st.max_queue, "graph_launches": st.launches,
                   "pool_gib": round(st.device_bytes / 2**30, 2), "arrival_rate": SERVE_RATE,
                   "sides": [2 ** L + 1 for L in SERVE_LEVELS]}
        log(f"[serve] {st.completed} solves in {st.wall_s:.2f} s ({st.solves_per_s:.1f}/s), "
            f"p50 {st.p50_ms:.1f} ms p99 {st.p99_ms:.1f} ms, max queue {st.max_queue}")
        if not args.no_fp64:
            # the plain FP64 GPU solve, one request at a time, on the first requests of the same stream
            sl, sk, sb = [], [], []
            for (_, L, cs, bs) in reqs[:SERVE_FP64_SAMPLE]:
                side = 2 ** L + 1
                k = H.DeviceArray((side - 1) ** 2, np.float64)
                b = H.DeviceArray(side * side, np.float64)
                H.fill_lognormal(k, 1.0, cs)
                H.fill_uniform(b, -1.0, 1.0, bs)
                sl.append(L)
                sk.append(k)
                sb.append(b)
            H.synchronize()
            tot, nd, cyc, bad = fp64_single_solves(sl, sk, sb, limit_s=20.0)
            del sk, sb
            serving["fp64_single_gpu_same_stream"] = {
                "solves_per_s": round(nd / tot, 3), "requests": nd, "seconds": round(tot, 3),
                "iterations_median": int(np.median(cyc)), "not_converged": bad,
                "what": "the first requests of the stream solved one at a time by honei_hip_mg_solve_pcg in FP64"}
            serving["speedup_vs_fp64_single_gpu"] = round(st.solves_per_s / (nd / tot), 2)
            log(f"[serve] FP64 single-system GPU solve on {nd} of the requests: {nd / tot:.1f}/s")
        if args.capacity_seconds > 0:
            serving["capacity"] = serving_capacity(args.capacity_seconds)
        if args.payload_seconds > 0:
            serving["host_payload"] = serving_payload(args.payload_seconds)

    # ---- single large system (2049^2, constant coefficient) -----------------------------
    xt = uniform_pm1(3, n)
    db = H.DeviceArray(n, np.float64)
    dxt = H.DeviceArray.from_host(xt)
    A64 = device_q1(SIDE, H.F64)
    H.call("honei_hip_product_smell_dv_double", dxt.ptr, db.ptr, A64.Aj.ptr, A64.Ax.ptr, A64.Arl.ptr, 0, n, 9,
           A64.stride, 0, 1, None)
    del A64, dxt
    def single_solve(inner_vt, tol):
        """Median time, cycles, final defect (relative for MIXED, absolute
        for the FP64 FIXED driver) and max error vs x_true on the resident
        2049^2 hierarchy: MIXED (one cycle per FP64 outer step) for an FP32
        inner hierarchy, pure FP64 multigrid (FIXED, cycles until the
        defect falls below tol times the initial one) for an FP64 one."""
        mg = H.Multigrid(inner_vt, 1, LEVEL)
        if inner_vt == H.F64:
            mg.configure(H.mg_config(n_max_iter=100, tolerance=tol))
        else:
            mg.configure(H.mg_config(n_max_iter=1, tolerance=tol, zero_inner_start=True, initial_zero=True))
        mg.build_q1(0.7)
        dx = H.DeviceArray(n, np.float64)

        def run():
            dx.zero()
            if inner_vt == H.F64:
                it, fd = mg.solve_fixed(db, dx)
                return it, fd
            outer, _inner, rd = mg.solve_mixed(db, dx)
            return outer, rd

        run()
        ts = []
        for _ in range(5):
            H.synchronize()
            t0 = time.perf_counter()
            cyc, rd = run()
            ts.append(time.perf_counter() - t0)
        e = float(np.abs(dx.get() - xt).max())
        del mg, dx
        return float(np.median(ts)), cyc, rd, e

    t1m, outer, rd, err = single_solve(H.F32, 1e-8)
    c3 = {}
    for key, vt in (("mixed_fp32_inner", H.F32), ("fp64_mg", H.F64)):
        tm, oc, r, e = single_solve(vt, 1e-10)
        c3[key] = {"solves_per_s": round(1.0 / tm, 3), "ms": round(tm * 1e3, 3), "cycles": oc,
                   ("final_defect" if vt == H.F64 else "rel_residual"): r, "max_err_vs_x_true": e}
        log(f"[2049] {key} to 1e-10: {tm * 1e3:.2f} ms, {oc} cycles, defect {r:.2e}, err {e:.2e}")
    c3["speedup_mixed_vs_fp64"] = round(c3["mixed_fp32_inner"]["solves_per_s"] / c3["fp64_mg"]["solves_per_s"], 2)
    c3["what"] = ("Q1 2049^2 (11 grids, 3x3 coarsest), b = A x_true (x_true ~ U[-1,1), seed 3), V(2,2) Jacobi 0.7, "
                  "CG on 3x3, relative defect 1e-10: MIXED (FP32 V-cycle, one per FP64 outer step) vs pure FP64 "
                  "multigrid (FIXED driver, FP64 V-cycles)")
    # the same system as a one-slot graph-launched pool (kappa = 1: the same operator)
    P1 = H.Batch([LEVEL], [1], H.batch_config(omega=0.7, iters_per_launch=8))
    ones = H.DeviceArray((SIDE - 1) ** 2, np.float64)
    ones.upload(np.ones(ones.n))
    xs = H.DeviceArray(n, np.float64)
    tp = []
    for k in range(6):
        H.synchronize()
        t0 = time.perf_counter()
        P1.load(0, ones, db)
        P1.run()
        tp.append(time.perf_counter() - t0)
    P1.result(0, xs)
    info1 = P1.poll()[0]
    err1 = float(np.abs(xs.get() - xt).max())
    del P1, xs, ones, db

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": ("synthetic: per-cell coefficient exp(N(0,1)) and b ~ U[-1,1), SplitMix64 seeds per system, "
                 "drawn on the device"),
        "config": {"workload": "batched_q1_lognormal_256_systems_sides_33_to_513_mixed_mg_to_1e-8",
                   "systems": len(levels),
                   "sides": {str(2 ** L + 1): levels.count(L) for L in sorted(set(levels))},
                   "unknowns": int(sum((2 ** L + 1) ** 2 for L in levels)),
                   "inner": "FP32 V(2,2) damped Jacobi 0.8, coefficient coarsened by geometric mean, CG on 3x3",
                   "outer": "FP64 conjugate gradients preconditioned by one V-cycle per iteration",
                   "rel_residual": 1e-8,
                   "step": "assemble all hierarchies from resident coefficients + solve all systems"},
        # the metric's latency is a request's: arrival to result in the
        # serving stream (every system of a batched step completes with it)
        "latency_ms": ({"p50": serving["latency_ms"]["p50"], "p99": serving["latency_ms"]["p99"],
                        "max": serving["latency_ms"]["max"],
                        "what": ("request latency, arrival to result, in the serving stream (Poisson arrivals "
                                 "at %g/s, sides 65..1025, to 1e-8)" % serving["arrival_rate"])}
                       if serving else None),
        "batch_completion_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 3),
                                "p99": round(float(np.percentile(lat_ms, 99)), 3),
                                "max": round(float(lat_ms.max()), 3),
                                "what": ("per-system completion time within a batched step (one graph launch: "
                                         "every system completes with the step)")},
        "solve": {"outer_iterations_median": int(np.median(iters)), "outer_iterations_max": int(iters.max()),
                  "max_rel_residual": float(rels.max()), "all_converged": bool((status == 0).all()),
                  "graph_launches_per_step": round(launches / args.steps, 2)},
        "bf16_inner": bf16,
        "fp64_batched": fp64b,
        "speedup_mixed_vs_fp64_batched": round(value / fp64b["solves_per_s"], 2),
        "fp64_single_gpu": fp64,
        "speedup_vs_fp64_single_gpu": round(value / fp64["solves_per_s"], 2) if fp64 else None,
        "serving": serving,
        "cg_q1_129_fp64": cg129,
        "pcg_jacobi_2049_fp64": pcg2049,
        "mg_q1_2049_to_1e-10": c3,
        "single_2049": {"solves_per_s": round(1.0 / t1m, 3), "outer_iterations": outer,
                        "rel_residual": rd, "max_err_vs_x_true": err,
                        "what": "one Q1 2049^2 constant-coefficient system, b = A x_true, mixed MG to 1e-8",
                        "as_one_slot_pool": {"solves_per_s": round(1.0 / float(np.median(tp[1:])), 3),
                                             "outer_iterations": info1.iters, "rel_residual": info1.rel,
                                             "max_err_vs_x_true": err1,
                                             "what": "same solve through a one-slot batch (graph launches, "
                                                     "device-side convergence; includes the hierarchy "
                                                     "assembly from the coefficient)"}},
        "spmv": spmv,
        "batch_kernel": batch_kernel,
        "step_roofline": step_rl,
    }
    r = spmv["q1_2049_f32"]
    out["roofline"] = {"bound": "hbm", "achieved": r["achieved_gbs"], "peak": PEAK_GBS, "unit": "GB/s",
                       "frac": round(r["achieved_gbs"] / PEAK_GBS, 4), "traffic": read_traffic(),
                       "kernel": "ell_row_kernel<float,1,ELL_PRODUCT> on the Q1 2049^2 operator, FP32",
                       "algorithmic_bytes_per_launch": r["algorithmic_bytes"]}
    if args.no_cpu:
        out["cpu_baseline"] = None
    else:
        one = cpu_baseline_batched(levels)
        out["cpu_baseline"] = cpu_all
        out["cpu_baseline_one_core"] = one
        # final error of the GPU results against the oracle's FP64 direct solve
        errs = fp64_oracle_errors(err_samples, cpu_workers())
        for key, obj in (("batch", out["solve"]), ("bf16_inner", bf16), ("fp64_batched", fp64b),
                         ("serving", serving)):
            if obj is not None and key in errs:
                obj["max_err_vs_fp64_oracle"] = errs[key]["max_err_vs_fp64_oracle"]
                obj["fp64_oracle_check"] = errs[key]
        out["fp64_oracle_check_what"] = (
            f"first {ERR_SAMPLE} systems of each size: GPU solution vs the oracle's FP64 direct solve of the same "
            "system (sparse LU + refinement to 1e-14 relative residual), max |x - x64| / max |x64|, and the "
            "oracle's FP64 relative residual of the GPU solution")
        if ref_proc is not None:
            p, q = ref_proc
            out["cpu_reference_path"] = q.get()
            p.join()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
