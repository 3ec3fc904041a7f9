"""CPU model of the paired row passes of the batched V-cycle (k_b_pair in
honei_amd/csrc/batch_kernels.hip): its tile geometry, the LDS staging of the
first pass's operand with a two-node halo, the first-pass tile and the
second pass read from it, restated in numpy float32 with the kernel's index
arithmetic and operation order.  Checked against two separate passes (the
T = 1 row sums of the separate launches): bit-identical results, every
global index in bounds, every output node written exactly once, and every
staged entry a pass actually uses belonging to the neighbour it stands for.
The GPU test test_paired_passes_equal_separate_launches checks the kernel
itself; this one pins the indexing on the CPU for sides the GPU tests do not
cover (including sides that are not 2^l + 1), for both tile heights: 16
staged rows (one grid row per thread, 14-row tiles) and 32 (two grid rows per
thread, tile row I on thread row I mod 16: 30-row tiles).  The thread that
computes a tile row does not enter the arithmetic, so the model indexes by
tile row.
"""
import numpy as np
import pytest

f32 = np.float32
KPAIR_WIDTH, KPAIR_TILE_ROWS = 62, 16   # kPairWidth, kPairTileRows (batch.hh)


def q1_level(S, rng):
    n = S * S
    stride = 32 * ((n + 31) // 32)
    Ax = np.zeros(9 * stride, f32)
    for r in range(S):
        for c in range(S):
            node = r * S + c
            if r in (0, S - 1) or c in (0, S - 1):
                Ax[node] = f32(1.0)
            else:
                w = rng.uniform(-0.5, -0.1, 9).astype(f32); w[4] = f32(rng.uniform(2, 4))
                for k in range(9):
                    Ax[node + k * stride] = w[k]
    return Ax, stride

def row_sum(Ax, stride, S, node, xv_of):
    r, c = divmod(node, S)
    bnd = r in (0, S - 1) or c in (0, S - 1)
    s = f32(0)
    for k in range(1 if bnd else 9):
        j = node if bnd else node + (k // 3 - 1) * S + (k % 3 - 1)
        s = f32(s + f32(Ax[node + k * stride] * xv_of(j)))
    return s

def two_pass(Ax, stride, S, rhs, dinv, xin, pair):
    n = S * S
    if pair == "pre":
        x = np.zeros(n, f32)
        for i in range(n):
            s = row_sum(Ax, stride, S, i, lambda j: f32(dinv[j] * rhs[j]))
            x[i] = f32(f32(dinv[i] * rhs[i]) + f32(f32(rhs[i] - s) * dinv[i]))
        d = np.zeros(n, f32)
        for i in range(n):
            s = row_sum(Ax, stride, S, i, lambda j: x[j])
            d[i] = f32(rhs[i] - s)
        return x, d
    x1 = np.zeros(n, f32)
    for i in range(n):
        s = row_sum(Ax, stride, S, i, lambda j: xin[j])
        x1[i] = f32(xin[i] + f32(f32(rhs[i] - s) * dinv[i]))
    x2 = np.zeros(n, f32)
    for i in range(n):
        s = row_sum(Ax, stride, S, i, lambda j: x1[j])
        x2[i] = f32(x1[i] + f32(f32(rhs[i] - s) * dinv[i]))
    return x2, None

def pair_kernel(Ax, stride, S, rhs, dinv, xin, pair, W=KPAIR_WIDTH, TR=KPAIR_TILE_ROWS):
    n = S * S
    R = TR - 2
    strips = (S + W - 1) // W; width = (S + strips - 1) // strips
    segs = (S + R - 1) // R
    assert width <= 62
    out0 = np.full(n, np.nan, f32); out1 = np.full(n, np.nan, f32) if pair == "pre" else None
    written0 = np.zeros(n, int); written1 = np.zeros(n, int)
    def chk(i, lim):
        assert 0 <= i < lim, (i, lim)
        return i
    for strip in range(strips):
        for seg in range(segs):
            tile = np.full((TR, 64), np.nan, f32)
            r0 = seg * R; r1 = min(S, r0 + R)
            c0 = strip * width; c1 = min(S, c0 + width)
            keep = {}
            def operand(gr, gc):
                gr = min(max(gr, 0), S - 1); gc = min(max(gc, 0), S - 1)
                jj = gr * S + gc
                return f32(dinv[jj] * rhs[jj]) if pair == "pre" else xin[jj]
            src = np.full((TR + 2, 66), np.nan, f32)
            srcpos = {}
            for i in range(TR):
                for lane in range(64):
                    src[i + 1][lane + 1] = operand(r0 - 1 + i, c0 - 1 + lane)
                    srcpos[(i + 1, lane + 1)] = (r0 - 1 + i, c0 - 1 + lane)
            for t in range(2 * 64 + 2 * (TR + 2)):
                if t < 64:
                    a, b = 0, t + 1
                elif t < 128:
                    a, b = TR + 1, t - 64 + 1
                else:
                    a = (t - 128) // 2; b = 65 if (t - 128) % 2 else 0
                src[a][b] = operand(r0 - 2 + a, c0 - 2 + b)
                srcpos[(a, b)] = (r0 - 2 + a, c0 - 2 + b)
            assert not np.isnan(src).any()
            for i in range(TR):
                r = r0 - 1 + i
                rc = min(max(r, 0), S - 1)
                for lane in range(64):
                    col = c0 - 1 + lane
                    in1 = 0 <= col < S and col <= c1
                    in2 = lane >= 1 and col < c1
                    colc = col if in1 else (0 if col < 0 else S - 1)
                    node = rc * S + colc
                    bnd = rc in (0, S - 1) or colc in (0, S - 1)
                    av = np.array([Ax[chk(node + k * stride, 9 * stride)] for k in range(9)], f32)
                    rh = rhs[chk(node, n)]; di = dinv[node]
                    xv = []
                    for k in range(9):
                        a = i + 1 if bnd else i + 1 + (k // 3 - 1)
                        b = lane + 1 if bnd else lane + 1 + (k % 3 - 1)
                        assert 0 <= a < TR + 2 and 0 <= b < 66
                        xv.append(src[a][b])
                        if in1 and 0 <= r < S and k < (1 if bnd else 9):
                            gr, gc = srcpos[(a, b)]
                            jj = node if bnd else node + (k // 3 - 1) * S + (k % 3 - 1)
                            assert (gr, gc) == divmod(jj, S), ((gr, gc), divmod(jj, S))
                    s_ = f32(0)
                    for k in range(1 if bnd else 9):
                        s_ = f32(s_ + f32(av[k] * xv[k]))
                    x0 = f32(di * rh) if pair == "pre" else xin[node]
                    v = f32(x0 + f32(f32(rh - s_) * di))
                    tile[i][lane] = v
                    if pair == "pre" and in2 and i >= 1 and r < r1:
                        out0[node] = v; written0[node] += 1
                    keep[(i, lane)] = (av, rh, di, bnd)
            for i in range(TR):
                q = r0 - 1 + i
                if not (1 <= i <= R and q < r1):
                    continue
                for lane in range(64):
                    col = c0 - 1 + lane
                    in2 = lane >= 1 and col < c1
                    if not in2:
                        continue
                    av, rh, di, bnd = keep[(i, lane)]
                    node = q * S + col
                    xv = []
                    for k in range(9):
                        ik = i if bnd else i + (k // 3 - 1)
                        lk = lane if bnd else lane + (k % 3 - 1)
                        assert 0 <= ik < TR and 0 <= lk < 64
                        val = tile[ik][lk]
                        if k < (1 if bnd else 9):
                            rr = r0 - 1 + ik; cc = c0 - 1 + lk
                            assert 0 <= rr < S and 0 <= cc < S, (rr, cc)
                        xv.append(val)
                    s_ = f32(0)
                    for k in range(1 if bnd else 9):
                        s_ = f32(s_ + f32(av[k] * xv[k]))
                    if pair == "pre":
                        out1[node] = f32(rh - s_); written1[node] += 1
                    else:
                        out0[node] = f32(tile[i][lane] + f32(f32(rh - s_) * di)); written0[node] += 1
    assert (written0 == 1).all(), np.unique(written0)
    if pair == "pre":
        assert (written1 == 1).all()
    return out0, out1


@pytest.mark.parametrize("side,tr", [(3, 16), (5, 16), (17, 16), (33, 16), (40, 16), (65, 16), (97, 16),
                                     (5, 32), (33, 32), (40, 32), (65, 32), (97, 32)])
@pytest.mark.parametrize("pair", ["pre", "post"])
def test_pair_tiles_equal_two_passes(side, tr, pair):
    rng = np.random.default_rng(side)
    Ax, stride = q1_level(side, rng)
    n = side * side
    rhs = rng.uniform(-1, 1, n).astype(f32)
    dinv = rng.uniform(0.2, 0.4, n).astype(f32)
    xin = rng.uniform(-1, 1, n).astype(f32)
    a0, a1 = two_pass(Ax, stride, side, rhs, dinv, xin, pair)
    b0, b1 = pair_kernel(Ax, stride, side, rhs, dinv, xin, pair, TR=tr)
    assert np.array_equal(a0.view(np.uint32), b0.view(np.uint32))
    if a1 is not None:
        assert np.array_equal(a1.view(np.uint32), b1.view(np.uint32))


# ---- logical-block enumeration of the launch (b_pair / k_b_pair) ----------
KXCDS = 8                      # kXcds (hip_common.hh)
KRESIDENT_THREADS = 256 * 2048  # kResidentThreads (batch_kernels.hip)


def xcd_block(b, nblocks):
    """hip_common.hh xcd_block: work-group b -> logical block."""
    q, r = divmod(nblocks, KXCDS)
    x, i = b % KXCDS, b // KXCDS
    return x * q + min(x, r) + i


KTILE_SLOTS = 1024            # kTileSlots (batch_kernels.hip)


def batch_grid(items, block):
    cap = KRESIDENT_THREADS // block
    return 1 if items < 1 else min(items, cap)


def tiled_grid(units, block):
    """batch_kernels.hip tiled_grid: enough work-groups that none walks more
    than kTileSlots logical blocks."""
    return max(batch_grid(units, block), (units + KTILE_SLOTS - 1) // KTILE_SLOTS)


def pair_geometry(S, W=KPAIR_WIDTH, R=KPAIR_TILE_ROWS - 2):
    strips = (S + W - 1) // W
    width = (S + strips - 1) // strips
    segs = (S + R - 1) // R
    return strips, width, segs


@pytest.mark.parametrize("nr", [1, 2])
@pytest.mark.parametrize("side,slots", [(33, 61), (65, 3), (257, 53), (513, 1), (513, 51), (1025, 1), (1025, 7),
                                        (1025, 300)])
def test_pair_launch_visits_every_tile_once(side, slots, nr):
    """The launch's grid of 1024-thread work-groups, each walking logical
    blocks lb = xcd_block(b) + k * grid, with lb -> (listed slot e, seg,
    strip) by the kernel's two divisions (strips minor): every tile of every
    listed slot is visited exactly once, no work-group walks more than the
    kTileSlots logical blocks whose slots it stages in LDS, consecutive
    logical blocks are the side-by-side tiles of one band of rows, and the
    tiles of a slot cover its grid exactly once with widths <= 62.  nr: grid
    rows per thread (tiles of 16 nr - 2 rows)."""
    R = KPAIR_TILE_ROWS * nr - 2
    strips, width, segs = pair_geometry(side, R=R)
    assert width <= KPAIR_WIDTH and (strips - 1) * width < side
    ups = strips * segs
    total = slots * ups
    grid = tiled_grid(total, KPAIR_TILE_ROWS * 64)
    seen = np.zeros((slots, segs, strips), np.int32)
    first_lb = {}
    for b in range(grid):
        lb = xcd_block(b, grid)
        mine = (total - 1 - lb) // grid + 1 if lb < total else 0
        assert mine <= KTILE_SLOTS
        while lb < total:
            e, rem = divmod(lb, ups)
            seg, strip = divmod(rem, strips)
            seen[e, seg, strip] += 1
            first_lb[(e, seg, strip)] = lb
            lb += grid
    assert (seen == 1).all()
    # strips minor: tile (seg, strip + 1) is the next logical block after (seg, strip)
    for (e, seg, strip), lb in list(first_lb.items())[:2000]:
        if strip + 1 < strips:
            assert first_lb[(e, seg, strip + 1)] == lb + 1
    cover = np.zeros((side, side), np.int32)
    for seg in range(segs):
        for strip in range(strips):
            r0, c0 = seg * R, strip * width
            cover[r0:min(side, r0 + R), c0:min(side, c0 + width)] += 1
    assert (cover == 1).all()
