// batch_kernels.hip -- kernels of the batched multigrid solver.
//
// Every batched launch walks the active slots list[a], a in
// [bounds[lo], bounds[hi]) (written by k_b_compact), on a fixed grid: the
// launch's logical blocks (row blocks of the active slots, slot-major) are
// split into contiguous chunks, one per work-group, with the XCD-aware
// work-group order, so the row blocks of one slot stay on one XCD's L2 and
// idle slots cost nothing.
//
// Row sums, transfers and the coarse CG use the same device code as the
// single-system kernels (ell_device.hh, transfer_device.hh,
// solver_device.hh), so a slot computes what the single-system solver
// computes.  The FP64 outer defect+norm also reproduces the single-system
// reduction order: 256-row blocks, partial b of a slot holding logical row
// block xcd_block(b, blocks per slot), partials summed by one 1024-thread
// block.

#include "batch.hh"
#include "ell_device.hh"
#include "q1_device.hh"
#include "solver_device.hh"
#include "transfer_device.hh"

#pragma clang fp contract(off)

namespace honei {
namespace hipx {
namespace {

// Launch grids hold about as many work-groups as the chip keeps resident
// (256 CUs x 2048 threads); work-group p walks logical blocks p, p + grid,
// p + 2 grid, ... with p = xcd_block(blockIdx.x, grid), so in each round the
// grid covers one contiguous window of logical blocks and every XCD a
// contiguous piece of it (few concurrent memory streams, neighbouring rows in
// one L2).  Logical block counts stay below 2^31 (checked by the launchers),
// so the per-block index arithmetic is 32-bit.
constexpr int64_t kResidentThreads = 256 * 2048;
constexpr int64_t kMaxLogicalBlocks = (int64_t(1) << 31) - 1;

__device__ inline int64_t first_block() { return xcd_block(blockIdx.x, gridDim.x); }

// q = a / b, a - q b for 0 <= a < 2^31, 0 < b < 2^31
__device__ inline void div32(int64_t a, int64_t b, int64_t &q, int64_t &r) {
    const uint32_t qq = (uint32_t)a / (uint32_t)b;
    q = qq;
    r = a - (int64_t)qq * b;
}

__device__ inline int64_t active_first(const BWork &w) { return w.bounds[w.lo]; }
__device__ inline int64_t active_count(const BWork &w) { return w.bounds[w.hi] - w.bounds[w.lo]; }

// list of non-idle slots (ascending) and, for each boundary k, the number of
// them below boundaries[k]; one 1024-thread block.
__device__ inline bool listed(const SlotScalars &s, int mode) {
    if (mode == LIST_ACTIVE) return s.state == SLOT_ACTIVE;
    if (mode == LIST_UNSETTLED) return s.state == SLOT_NEW || s.state == SLOT_VERIFY;
    return s.state != SLOT_IDLE;
}

__global__ __launch_bounds__(1024) void k_b_compact(const SlotScalars *sc, int64_t nslots, BBoundaries bd,
                                                     int32_t *list, int32_t *bounds, int mode) {
    __shared__ int32_t pre[1024];
    const int t = threadIdx.x;
    const int64_t chunk = (nslots + 1023) / 1024;
    const int64_t s0 = t * chunk, s1 = (s0 + chunk < nslots) ? s0 + chunk : nslots;
    int32_t cnt = 0;
    for (int64_t g = s0; g < s1; ++g) cnt += listed(sc[g], mode);
    // inclusive scan over the 1024 counts: within each wavefront, then over
    // the wavefront totals (integer sums: any order gives the same counts)
    __shared__ int32_t wtot[1024 / kWave];
    const int lane = t % kWave, wv = t / kWave;
    int32_t v = cnt;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int32_t u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    if (lane == kWave - 1) wtot[wv] = v;
    __syncthreads();
    if (wv == 0) {
        int32_t x = lane < 1024 / kWave ? wtot[lane] : 0;
#pragma unroll
        for (int off = 1; off < 1024 / kWave; off <<= 1) {
            const int32_t u = __shfl_up(x, off);
            if (lane >= off) x += u;
        }
        if (lane < 1024 / kWave) wtot[lane] = x;
    }
    __syncthreads();
    pre[t] = v + (wv > 0 ? wtot[wv - 1] : 0);
    __syncthreads();
    int32_t pos = pre[t] - cnt;
    for (int64_t g = s0; g < s1; ++g)
        if (listed(sc[g], mode)) list[pos++] = (int32_t)g;
    if (t < bd.n) {
        const int64_t b = bd.v[t];
        const int64_t owner = chunk > 0 ? b / chunk : 0;
        int32_t c = (owner > 0 && owner <= 1023) ? pre[owner - 1] : (owner > 1023 ? pre[1023] : 0);
        if (owner <= 1023)
            for (int64_t g = owner * chunk; g < b && g < nslots; ++g) c += listed(sc[g], mode);
        bounds[t] = c;
    }
}

// One logical block = blockDim/T rows of a group of G consecutive active
// slots, all of one level: each index is loaded once for the group, each
// slot's sums are the single-slot sums.  A short last group repeats its first
// slot in the spare places and stores only its own slots.
template <typename VT, int T, int MODE, int G>
__global__ __launch_bounds__(1024) void k_b_ell(BMat m, BWork w, void *out, const void *x, const void *rhs,
                                                const void *dinv) {
    typedef typename VecOf<VT>::type V;
    const int64_t a0 = active_first(w), na = active_count(w);
    const int64_t total = ((na + G - 1) / G) * m.bps;
    for (int64_t lb = first_block(); lb < total; lb += gridDim.x) {
        int64_t grp, rb_now;
        div32(lb, m.bps, grp, rb_now);
        const int64_t e0 = grp * G;
        const int cnt = (int)((na - e0) < G ? (na - e0) : G);
        const VT *ax[G];
        const V *xs[G], *rs[G], *ds[G];
        V *os[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int64_t ls = w.list[a0 + e0 + (g < cnt ? g : 0)] - w.base;
            const int64_t vo = ls * m.vld;
            ax[g] = (const VT *)m.Ax + ls * m.ax_ld;
            xs[g] = (const V *)x + vo;
            rs[g] = (const V *)rhs + vo;
            ds[g] = (const V *)dinv + vo;
            os[g] = (V *)out + vo;
        }
        const int64_t gid = rb_now * (int64_t)blockDim.x + threadIdx.x;
        const int lane = (int)(gid & (T - 1));
        const int64_t row = gid / T;
        const bool active = row < m.n;
        V s[G];
        if (active) {
            // batch levels are Q1 operators: stored width q1_width(T) by construction
            if constexpr (MODE == ELL_JACOBI0)
                ell_lane_sum_group_w<VT, T, G, q1_width(T), true>(m.Aj, ax, row, lane, m.Arl[row], m.stride, rs, ds,
                                                                  s);
            else
                ell_lane_sum_group_w<VT, T, G, q1_width(T), false>(m.Aj, ax, row, lane, m.Arl[row], m.stride, xs, ds,
                                                                   s);
        } else {
#pragma unroll
            for (int g = 0; g < G; ++g) s[g] = V(0);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) s[g] = lane_group_sum<T>(s[g]);
        if (active && lane == 0) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (g < cnt) {
                    if constexpr (MODE == ELL_DEFECT) {
                        os[g][row] = rs[g][row] - s[g];
                    } else {
                        const V xin = (MODE == ELL_JACOBI0) ? ds[g][row] * rs[g][row] : xs[g][row];
                        const V t = rs[g][row] - s[g];
                        os[g][row] = xin + t * ds[g][row];
                    }
                }
            }
        }
    }
}

// ---- two dependent row passes of a Q1 level in one launch --------------------
// The levels of the batch are lexicographic side x side Q1 grids: an
// interior row holds its 9-point stencil in column order (neighbour k at
// grid offset (k/3 - 1, k%3 - 1)), a boundary row its diagonal alone
// (q1_row), T = 1 layout.  k_b_pair runs two passes that depend on each
// other in one launch.  A work-group of kPairTileRows wavefronts owns a tile
// of one slot's grid: kPairTileRows * NR - 2 rows (NR = 1 or 2 grid rows per
// thread) by `width` (<= 62) columns.  The work-group first stages the first
// pass's gathered operand for the tile and a two-node halo in LDS; after a
// barrier, tile row I (thread row I mod kPairTileRows) forms the first pass
// for grid row r0 - 1 + I on the tile's columns plus one halo column each
// side (a lane per column) from it into LDS; after a second barrier, tile
// rows 1 .. kPairTileRows * NR - 2 form the second pass for their row from LDS,
// with the matrix row still in registers.  So the matrix is read from
// HBM once for both passes: halo rows and columns are read again by the
// neighbouring tiles, through L2 (cached loads, not streaming ones).  Every
// node's sum is the T = 1 sum of k_b_ell (same terms, same order, same
// epilogue), so the result is bit-identical to two separate launches: an
// interior node's nine terms in column order, a boundary node's 0 + 1 * x
// (selected after every node has formed the interior sum over its staged
// neighbours, so the LDS reads take fixed offsets).
//   PAIR_PRE : x = Jacobi0 (two sweeps from zero) -> out0, then the defect
//              d = rhs - A x -> out1;
//   PAIR_POST: out0 = J(J(xin)) (xin != out0); with FP (the prolongation
//              fused in) out0 = J(J(xin + P xc)), the operand xin + P xc
//              formed while staging with the terms and order of
//              k_b_prolong_rows (so the same bits as that launch followed
//              by the unfused pass pair).
struct PairGeom {
    int32_t side, width, strips, segs;
};

// Tiled launches stage the slots of a work-group's tiles in LDS at the
// start; their grids give no work-group more than kTileSlots tiles.
constexpr int kTileSlots = 1024;
// The element-wise outer launches stage at most kElemSlots logical blocks'
// slots and scalars per work-group.
constexpr int kElemSlots = 64;
// The fused coarse cycle's work-group: one per slot, five wavefronts, so a
// sweep of the largest fused level (17^2 = 289 nodes) is one round.
constexpr int kFusedThreads = 320;
inline int tiled_grid(int64_t units, int threads) {
    const int64_t g = batch_grid(units, threads), need = (units + kTileSlots - 1) / kTileSlots;
    return (int)(g > need ? g : need);
}

// The four coarse values around fine node (R, C) of a (2 m + 1)^2 level
// (coarse side nc = m + 1): (R/2, C/2), (R/2, (C+1)/2), ((R+1)/2, C/2),
// ((R+1)/2, (C+1)/2), rounded down.
template <typename V>
__device__ inline void prolong_taps(const V *xc, int32_t nc, int32_t R, int32_t C, V (&t)[4]) {
    const int32_t rlo = R >> 1, rhi = (R + 1) >> 1, clo = C >> 1, chi = (C + 1) >> 1;
    // the two rows' values at columns c0, c0 + 1 (two adjacent loads each,
    // merged into one), c0 = clo except on the last column (nc >= 2)
    const int32_t c0 = clo < nc - 1 ? clo : nc - 2;
    const V *p0 = xc + rlo * nc + c0, *p1 = xc + rhi * nc + c0;
    const V a0 = p0[0], a1 = p0[1], b0 = p1[0], b1 = p1[1];
    t[0] = clo == c0 ? a0 : a1;
    t[1] = chi == c0 ? a0 : a1;
    t[2] = clo == c0 ? b0 : b1;
    t[3] = chi == c0 ? b0 : b1;
}
// prolong_node's correction of fine node (R, C) of an nf x nf level from its
// taps, every border Dirichlet (0): a coarse node's value, the mean of two
// (odd row or odd column) or of four (both odd), summed in that order.
template <typename V>
__device__ inline V prolong_value(int32_t R, int32_t C, int32_t nf, const V (&t)[4]) {
    const V h = V(0.5), q = V(0.25);
    const V f_rc = ((V(0) + t[0] * q + t[1] * q) + t[2] * q) + t[3] * q;
    const V f_c = (V(0) + t[0] * h) + t[1] * h;
    const V f_r = (V(0) + t[0] * h) + t[2] * h;
    const bool oddR = (R & 1) != 0, oddC = (C & 1) != 0;
    const V f = oddR ? (oddC ? f_rc : f_r) : (oddC ? f_c : t[0]);
    return (R == 0 || R == nf - 1 || C == 0 || C == nf - 1) ? V(0) : f;
}

template <typename VT> __device__ inline typename VecOf<VT>::type load_value_cached(const VT *p) { return *p; }
template <> __device__ inline float load_value_cached<bf16>(const bf16 *p) { return bf16_to_float(p->bits); }
// A node's four interleaved distinct weights in one (float, bf16) or two
// (double) vector loads; p is 4-entry aligned.
__device__ inline void load_weights4(const float *p, float (&o)[4]) {
    const float4 v = *(const float4 *)p;
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
}
__device__ inline void load_weights4(const bf16 *p, float (&o)[4]) {
    const uint2 v = *(const uint2 *)p;
    o[0] = bf16_to_float((uint16_t)(v.x & 0xffffu));
    o[1] = bf16_to_float((uint16_t)(v.x >> 16));
    o[2] = bf16_to_float((uint16_t)(v.y & 0xffffu));
    o[3] = bf16_to_float((uint16_t)(v.y >> 16));
}
__device__ inline void load_weights4(const double *p, double (&o)[4]) {
    const double2 a = *(const double2 *)p, b = *(const double2 *)(p + 2);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = b.x;
    o[3] = b.y;
}

// V: the vector and arithmetic type (float for float / bf16 values, double
// for double values)
// WM: weights from the stored values (0), formed from the cell coefficients
// (1), gathered from the distinct-weight arrays (2).
// FP (PAIR_POST only): xc = the coarse level's result (vld_c per slot).
// NR: grid rows per thread; thread (i, lane) takes tile rows i + u *
// kPairTileRows (u < NR), so a tile is kPairTileRows * NR - 2 rows and every
// thread keeps NR nodes' loads in flight.
template <typename VT, int PAIR, int WM, bool FP = false, int NR = 1, typename V = typename VecOf<VT>::type>
__global__ __launch_bounds__(kPairTileRows * kWave, 8) void k_b_pair(BMat m, BWork w, PairGeom pg, V *out0, V *out1,
                                                                  const V *xin, const V *rhs, const V *dinv,
                                                                  const V *xc, int64_t vld_c) {
    static_assert(!FP || PAIR == PAIR_POST, "the prolongation fuses into the post-smoothing pair");
    static_assert(NR == 1 || NR == 2, "one or two rows per thread");
    constexpr int TR = kPairTileRows * NR, OR = TR - 2;   // staged rows, tile rows
    // first-pass results of the tile and its halo
    __shared__ V tile[TR][kWave];
    // the first pass's gathered operand (PAIR_PRE: dinv .* rhs; PAIR_POST:
    // xin) on grid rows r0 - 2 .. r0 + TR - 1 and columns c0 - 2 ..
    // c0 + kWave - 1 (entry [a][b] = grid (r0 - 2 + a, c0 - 2 + b), positions
    // outside the grid clamped to it: loaded, never used)
    __shared__ V src[TR + 2][kWave + 2];
    constexpr int kExtra = 2 * kWave + 2 * (TR + 2);   // src entries beyond one per node
    const int i = threadIdx.x / kWave, lane = threadIdx.x % kWave;   // tile row (u = 0), lane
    const int64_t a0 = active_first(w), na = active_count(w);
    const int64_t ups = (int64_t)pg.strips * pg.segs;
    const int64_t total = na * ups;
    const int S = pg.side;
    // the slots of this work-group's tiles (at most kTileSlots: the launcher
    // sizes the grid so), loaded at once: one list round trip per launch
    // instead of one per tile
    __shared__ int32_t tslot[kTileSlots];
    const int64_t lb0 = first_block();
    const int64_t mine = lb0 < total ? (total - 1 - lb0) / gridDim.x + 1 : 0;
    for (int64_t k = threadIdx.x; k < mine; k += blockDim.x) tslot[k] = w.list[a0 + (lb0 + k * gridDim.x) / ups];
    __syncthreads();
    // the staged entry beyond one per node (threads < kExtra): grid rows
    // r0 - 2 and r0 + TR - 1, columns c0 - 2 and c0 + kWave - 1
    int ea = 0, eb = 0;
    if (threadIdx.x < kExtra) {
        const int t = threadIdx.x;
        if (t < kWave) {
            ea = 0;
            eb = t + 1;
        } else if (t < 2 * kWave) {
            ea = TR + 1;
            eb = t - kWave + 1;
        } else {
            ea = (t - 2 * kWave) / 2;
            eb = ((t - 2 * kWave) % 2) ? kWave + 1 : 0;
        }
    }
    int64_t k = 0;
    for (int64_t lb = lb0; lb < total; lb += gridDim.x, ++k) {
        int64_t e, rem, strip, seg;
        div32(lb, ups, e, rem);
        // strips minor: the tiles side by side in a band of rows are
        // consecutive logical blocks, so they run together on one XCD and the
        // cache lines their rows share (a row of a tile starts at an arbitrary
        // column) are fetched once into that XCD's L2
        div32(rem, pg.strips, seg, strip);
        const int64_t ls = tslot[k] - w.base;
        const int64_t vo = ls * m.vld;
        const VT *ax = (const VT *)m.Ax + ls * m.ax_ld;
        const V *xs = xin + vo, *rs = rhs + vo, *ds = dinv + vo;
        const int r0 = (int)seg * OR, r1 = min(S, r0 + OR);
        const int c0 = (int)strip * pg.width, c1 = min(S, c0 + pg.width);
        const int col = c0 - 1 + lane;
        const bool in2 = lane >= 1 && col < c1;                    // second-pass lanes
        const int colc = col < 0 ? 0 : (col > S - 1 ? S - 1 : col);
        // this thread's nodes: grid (r[u], col) clamped into the grid (lanes
        // and rows outside it are computed and never used)
        int r[NR], rc[NR];
        int64_t node[NR];
        bool bnd[NR];
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            r[u] = r0 - 1 + i + u * kPairTileRows;
            rc[u] = r[u] < 0 ? 0 : (r[u] > S - 1 ? S - 1 : r[u]);
            node[u] = (int64_t)rc[u] * S + colc;
            bnd[u] = rc[u] == 0 || rc[u] == S - 1 || colc == 0 || colc == S - 1;
        }
        const int egr = min(max(r0 - 2 + ea, 0), S - 1), egc = min(max(c0 - 2 + eb, 0), S - 1);
        const int64_t ej = (int64_t)egr * S + egc;
        // ---- every global load of the tile issued before any use ----
        V rh[NR], di[NR], xn[NR];
        V tn[NR][4], te[4];   // FP: coarse taps of the nodes and of the extra entry
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            rh[u] = rs[node[u]];
            di[u] = ds[node[u]];
            xn[u] = (PAIR == PAIR_PRE) ? V(0) : xs[node[u]];
            if constexpr (FP) prolong_taps(xc + ls * vld_c, (S >> 1) + 1, rc[u], colc, tn[u]);
        }
        // the extra entry's loads only where there is one (the first waves)
        V e1 = V(0), e2 = V(0);
        if (threadIdx.x < kExtra) {
            e1 = (PAIR == PAIR_PRE) ? ds[ej] : xs[ej];
            if (PAIR == PAIR_PRE) e2 = rs[ej];
            if constexpr (FP) prolong_taps(xc + ls * vld_c, (S >> 1) + 1, egr, egc, te);
        }
        V av[NR][9];
        double cw[NR][4];
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            if constexpr (WM == 1) {
                // the node's four cells (a boundary node's: clamped, unused)
                const uint32_t nc = (uint32_t)S - 1u;
                const uint32_t ir = (uint32_t)min(max(rc[u], 1), S - 2), ic = (uint32_t)min(max(colc, 1), S - 2);
                const double *cs = m.kappa + ls * m.kld + ((ir - 1u) * nc + (ic - 1u)), *cs2 = cs + nc;
                cw[u][0] = cs[0];
                cw[u][1] = cs[1];
                cw[u][2] = cs2[0];
                cw[u][3] = cs2[1];
            } else if constexpr (WM == 2) {
                // the interleaved distinct weights (cell, row pair, column
                // pair, diagonal) of the node and of its neighbours at -1, -S,
                // -S-1: cell weights at node - S - 1, node - S, node - 1,
                // node; row-pair weights at node - S, node; column-pair
                // weights at node - 1, node; the diagonal at node (boundary
                // rows: addresses of the node itself, values unused)
                const VT *sw4 = (const VT *)m.st + ls * 4 * m.vld;
                const int64_t nd = node[u], dS = bnd[u] ? 0 : S, d1 = bnd[u] ? 0 : 1;
                V q0[4], q1[4], q2[4];
                load_weights4(sw4 + 4 * nd, q0);
                load_weights4(sw4 + 4 * (nd - d1), q1);
                load_weights4(sw4 + 4 * (nd - dS), q2);
                av[u][0] = load_value_cached(sw4 + 4 * (nd - dS - d1));
                av[u][1] = q2[1];
                av[u][2] = q2[0];
                av[u][3] = q1[2];
                av[u][4] = q0[3];
                av[u][5] = q0[2];
                av[u][6] = q1[0];
                av[u][7] = q0[1];
                av[u][8] = q0[0];
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) av[u][q] = load_value_cached(ax + node[u] + (int64_t)q * m.stride);
            }
        }
        if constexpr (FP) {   // t = x + P xc, as k_b_prolong_rows stores it
#pragma unroll
            for (int u = 0; u < NR; ++u) xn[u] = xn[u] + prolong_value(rc[u], colc, S, tn[u]) * V(1);
            if (threadIdx.x < kExtra) e1 = e1 + prolong_value(egr, egc, S, te) * V(1);
        }
        // ---- stage the first pass's operand ----
#pragma unroll
        for (int u = 0; u < NR; ++u)
            src[i + u * kPairTileRows + 1][lane + 1] = (PAIR == PAIR_PRE) ? di[u] * rh[u] : xn[u];
        if (threadIdx.x < kExtra) src[ea][eb] = (PAIR == PAIR_PRE) ? e1 * e2 : e1;
        // ---- the weights of the nodes' rows (a boundary node's: those of
        // its clamped cells, unused) ----
        if constexpr (WM == 1) {
#pragma unroll
            for (int u = 0; u < NR; ++u)
                q1_interior_weights_cells<VT, V>(cw[u][0], cw[u][1], cw[u][2], cw[u][3], av[u]);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            // every node forms the interior stencil sum over its staged
            // neighbours; a boundary node then takes its identity row's sum
            // 0 + 1 * x instead
            const int I = i + u * kPairTileRows;
            V xv[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) xv[q] = src[I + 1 + (q / 3 - 1)][lane + 1 + (q % 3 - 1)];
            V sm = V(0);
#pragma unroll
            for (int q = 0; q < 9; ++q) sm = sm + av[u][q] * xv[q];
            if (bnd[u]) sm = V(0) + V(1) * xv[4];
            const V x0 = (PAIR == PAIR_PRE) ? di[u] * rh[u] : xn[u];
            const V t = rh[u] - sm;
            const V v1 = x0 + t * di[u];
            tile[I][lane] = v1;
            if (PAIR == PAIR_PRE && in2 && I >= 1 && r[u] < r1) out0[vo + node[u]] = v1;
        }
        __syncthreads();
        // ---- second pass, grid row r[u] for tile rows 1 .. OR ----
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            const int I = i + u * kPairTileRows;
            if (I >= 1 && I <= OR && r[u] < r1 && in2) {
                V xv[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) xv[q] = tile[I + (q / 3 - 1)][lane + (q % 3 - 1)];
                V sm = V(0);
#pragma unroll
                for (int q = 0; q < 9; ++q) sm = sm + av[u][q] * xv[q];
                if (bnd[u]) sm = V(0) + V(1) * xv[4];
                if constexpr (PAIR == PAIR_PRE) {
                    out1[vo + node[u]] = rh[u] - sm;
                } else {
                    const V t = rh[u] - sm;
                    out0[vo + node[u]] = tile[I][lane] + t * di[u];
                }
            }
        }
        __syncthreads();
    }
}

// ---- FP64 outer work, all classes in one launch ------------------------------
// Both outer kernels order their logical blocks class by class (per-class
// prefix in LDS, indexed by the class) and walk each slot's rows in order.
// Reductions store partials per slot (stride pst) in one of two layouts:
//   DC (the defect of defect correction): one partial per 256-row block,
//     partial index b holding row block xcd_block(b, bps) of that slot alone
//     summed by block_sum, the order a single-system launch of 256-row
//     blocks uses, so the batched solve reproduces the single-system norm
//     bit for bit;
//   otherwise (the CG reductions): one partial per 64-row chunk, the
//     wavefront sum of that chunk (nps = ceil(n / 64) partials per slot), so
//     no launch waits at a work-group barrier.

// Row sum of G slots' FP64 Q1 operators (as q1_row builds them, T = 1):
// the neighbour columns come from the grid position (interior row: the
// 9-point stencil in column order; boundary row: the diagonal alone), so no
// index array is read; the values, the guarded sum and its order are those of
// ell_lane_sum_group_w over the stored matrix, hence the same bits.
// MF: the values from the slots' cell coefficients (kap[g], q1_stored_weights)
// instead of Ax[g] -- the same doubles.
template <int G, bool MF>
__device__ inline void q1_stencil_sum_group(int64_t side, int64_t row, const double *const (&Ax)[G], int64_t stride,
                                            const double *const (&kap)[G], const double *const (&x)[G],
                                            double (&s)[G]) {
    const uint32_t r = (uint32_t)row / (uint32_t)side, c = (uint32_t)row - r * (uint32_t)side;
    const bool bnd = r == 0 || r == side - 1 || c == 0 || c == side - 1;
    int64_t j[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) j[k] = bnd ? row : row + (int64_t)(k / 3 - 1) * side + (k % 3 - 1);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        double av[9], xv[9];
        if constexpr (MF) q1_stored_weights<double, double>(side, r, c, bnd, kap[g], av);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if constexpr (!MF) av[k] = load_value_nt(Ax[g] + row + (int64_t)k * stride);
            xv[k] = x[g][j[k]];
        }
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) t = (k < (bnd ? 1 : 9)) ? t + av[k] * xv[k] : t;
        s[g] = t;
    }
}

// Defects b - A x: logical block = 256 rows of a group of G
// consecutive listed slots of one class (a short last group repeats its
// first slot in the spare places and stores only its own slots).  The
// neighbour columns come from the grid position, so slots share no loads;
// the launches use G = kOuterRowGroup.
template <int OP, int G, bool DC, bool MF>
__global__ __launch_bounds__(256) void k_b_outer_rows(BClassTab tab, BWork w, const SlotScalars *sc) {
    static_assert(OP == OP_DEFECT_ALL || OP == OP_DEFECT_SETTLE, "defect op (q = A p: k_b_dir_spmv)");
    __shared__ int64_t pre[kMaxClasses + 1];
    if (threadIdx.x == 0) {
        int64_t acc = 0;
        pre[0] = 0;
        for (int c = 0; c < tab.nclass; ++c) {
            acc += (int64_t)((w.bounds[c + 1] - w.bounds[c] + G - 1) / G) * tab.c[c].bps;
            pre[c + 1] = acc;
        }
    }
    __syncthreads();
    const int64_t total = pre[tab.nclass];
    for (int64_t lb = first_block(); lb < total; lb += gridDim.x) {
        int c = 0;
        while (lb >= pre[c + 1]) ++c;
        const BClassEntry &e = tab.c[c];
        int64_t gi, hb;   // group of the class list, row block
        div32(lb - pre[c], e.bps, gi, hb);
        const int64_t l0 = w.bounds[c] + gi * G;
        const int cnt = (int)((w.bounds[c + 1] - l0) < G ? (w.bounds[c + 1] - l0) : G);
        int64_t ls[G];
#pragma unroll
        for (int k = 0; k < G; ++k) ls[k] = w.list[l0 + (k < cnt ? k : 0)] - e.first;
        const int64_t row = hb * 256 + threadIdx.x;
        double part[G];
#pragma unroll
        for (int k = 0; k < G; ++k) part[k] = 0.0;
        if (row < e.n) {
            const double *ax[G], *xv[G], *kp[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                ax[k] = e.Ax + ls[k] * e.ax_ld;
                xv[k] = e.x + ls[k] * e.vld;
                kp[k] = MF ? e.kappa + (ls[k] + e.first) * e.kld : nullptr;
            }
            double sum[G];
            // the FP64 outer operator: Q1 (q1_row), one lane per row
            q1_stencil_sum_group<G, MF>(e.side, row, ax, e.stride, kp, xv, sum);
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const int64_t vi = ls[k] * e.vld + row;
                const double r = e.b[vi] - sum[k];
                if (k < cnt) e.r[vi] = r;
                part[k] = r * r;
            }
        }
        if constexpr (DC) {
            block_sum_first<G>(part);
            if (threadIdx.x == 0) {
                const int64_t pidx = xcd_unblock(hb, e.bps);
#pragma unroll
                for (int k = 0; k < G; ++k)
                    if (k < cnt) e.partials[ls[k] * e.pst + pidx] = part[k];
            }
        } else {
            const int64_t chunk = hb * (256 / kWave) + threadIdx.x / kWave;
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const double v = wave_sum(part[k]);
                if (threadIdx.x % kWave == 0 && k < cnt && chunk < e.nps) e.partials[ls[k] * e.pst + chunk] = v;
            }
        }
    }
}

// CG direction and q = A p in one tiled launch.  Logical block = one tile of
// one listed slot's finest grid (kDirRows rows by e.tw columns: the paired
// passes' column strips, kDirNodes grid rows per thread), with a one-node
// halo: every thread forms p = z + beta p_old at its nodes (z = |r| x_top;
// the node clamped into the grid) from the same two loads and expression
// wherever that node is formed, so the tiles agree on every p; the tile's own
// nodes store it into the slot's other direction buffer (the current one
// stays intact for the neighbouring tiles' halos).  After a barrier the own
// nodes form q = A p from LDS with the row sums of k_b_outer_rows (interior:
// the nine terms in column order; boundary: 0 + p), store q, and the tile
// stores one partial p.q (each thread's nodes in order, wave sums, then its
// wavefronts in order) at its tile index.  FIN_ALPHA sums the slot's
// e.ntiles partials and makes the new buffer current.  Z: the inner vectors'
// type.
template <bool MF, typename Z>
__global__ __launch_bounds__(kPairTileRows * kWave, 8) void k_b_dir_spmv(BClassTab tab, BWork w,
                                                                        const SlotScalars *sc) {
    constexpr int NR = kDirNodes, TR = kPairTileRows * NR;   // grid rows per thread, staged rows
    __shared__ int64_t pre[kMaxClasses + 1];
    __shared__ double sp[TR][kWave];
    __shared__ double red[kPairTileRows];
    if (threadIdx.x == 0) {
        int64_t acc = 0;
        pre[0] = 0;
        for (int c = 0; c < tab.nclass; ++c) {
            acc += (int64_t)(w.bounds[c + 1] - w.bounds[c]) * tab.c[c].ntiles;
            pre[c + 1] = acc;
        }
    }
    __syncthreads();
    const int i = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t total = pre[tab.nclass];
    // the slots and scalars of this work-group's tiles (at most kTileSlots:
    // the launcher sizes the grid so), loaded at once: two round trips per
    // launch instead of two per tile
    __shared__ int32_t tslot[kTileSlots];
    __shared__ double tbeta[kTileSlots], tscale[kTileSlots];
    __shared__ uint8_t tcur[kTileSlots];
    const int64_t lb0 = first_block();
    const int64_t mine = lb0 < total ? (total - 1 - lb0) / gridDim.x + 1 : 0;
    for (int64_t k = threadIdx.x; k < mine; k += blockDim.x) {
        const int64_t lb = lb0 + k * gridDim.x;
        int c = 0;
        while (lb >= pre[c + 1]) ++c;
        const int64_t g = w.list[w.bounds[c] + (lb - pre[c]) / tab.c[c].ntiles];
        const SlotScalars &st = sc[g];
        tslot[k] = (int32_t)g;
        tbeta[k] = st.beta;
        tscale[k] = st.scale;
        tcur[k] = (uint8_t)(st.pcur & 1);
    }
    __syncthreads();
    int64_t kt = 0;
    for (int64_t lb = lb0; lb < total; lb += gridDim.x, ++kt) {
        int c = 0;
        while (lb >= pre[c + 1]) ++c;
        const BClassEntry &e = tab.c[c];
        int64_t si, tile, seg, strip;   // listed slot of the class, tile of the slot
        div32(lb - pre[c], e.ntiles, si, tile);
        div32(tile, e.strips, seg, strip);
        const int64_t g = tslot[kt];
        const int64_t ls = g - e.first;
        const double beta = tbeta[kt], scale = tscale[kt];
        const int64_t cur = tcur[kt];
        const int S = (int)e.side;
        const int r0 = (int)seg * kDirRows, r1 = min(S, r0 + kDirRows);
        const int c0 = (int)strip * e.tw, c1 = min(S, c0 + e.tw);
        const int col = c0 - 1 + lane;
        const int colc = col < 0 ? 0 : (col > S - 1 ? S - 1 : col);
        const uint32_t ic = (uint32_t)min(max(colc, 1), S - 2);
        const double *pold = e.p + ls * e.vld + cur * e.palt;
        double *pnew = e.p + ls * e.vld + (1 - cur) * e.palt;
        const Z *zs = (const Z *)e.x_top + g * e.vld_top;
        int64_t node[NR];
        bool own[NR], bnd[NR];
        double zv[NR], po[NR], pv[NR];
        double av[NR][9];
        double cw[NR][4];
        // ---- every global load issued before any use ----
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            const int r = r0 - 1 + i + u * kPairTileRows;
            const int rc = r < 0 ? 0 : (r > S - 1 ? S - 1 : r);
            node[u] = (int64_t)rc * S + colc;
            own[u] = r >= r0 && r < r1 && lane >= 1 && col < c1;
            bnd[u] = rc == 0 || rc == S - 1 || colc == 0 || colc == S - 1;
            // interior position of the row's weights (a boundary node's: unused)
            const uint32_t ir = (uint32_t)min(max(rc, 1), S - 2);
            zv[u] = (double)zs[node[u]];
            po[u] = beta == 0.0 ? 0.0 : pold[node[u]];
            if constexpr (MF) {
                const uint32_t nc = (uint32_t)S - 1u;
                const double *cs = e.kappa + g * e.kld + ((ir - 1u) * nc + (ic - 1u)), *cs2 = cs + nc;
                cw[u][0] = cs[0];
                cw[u][1] = cs[1];
                cw[u][2] = cs2[0];
                cw[u][3] = cs2[1];
            } else {
                const double *ax = e.Ax + ls * e.ax_ld + ((int64_t)ir * S + ic);
#pragma unroll
                for (int k = 0; k < 9; ++k) av[u][k] = load_value_nt(ax + (int64_t)k * e.stride);
            }
        }
        // ---- p at the tile and its halo ----
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            const double z = zv[u] * scale;
            pv[u] = beta == 0.0 ? z : z + beta * po[u];
            sp[i + u * kPairTileRows][lane] = pv[u];
            if (own[u]) pnew[node[u]] = pv[u];
        }
        __syncthreads();
        // ---- q = A p at the own nodes, the tile's p.q ----
        double part = 0.0;
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            if constexpr (MF) q1_cell_weights<double>(cw[u][0], cw[u][1], cw[u][2], cw[u][3], av[u]);
            const int I = i + u * kPairTileRows;
            if (own[u]) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < 9; ++k) t = t + av[u][k] * sp[I + (k / 3 - 1)][lane + (k % 3 - 1)];
                if (bnd[u]) t = 0.0 + pv[u];
                e.q[ls * e.vld + node[u]] = t;
                part = part + pv[u] * t;
            }
        }
        const double v = wave_sum(part);
        if (lane == 0) red[i] = v;
        __syncthreads();
        if (i == 0) {
            const double t = wave_sum(lane < kPairTileRows ? red[lane] : 0.0);
            if (lane == 0) e.partials[ls * e.pst + tile] = t;
        }
    }
}

// Element-wise ops: logical block = R consecutive 256-row blocks of one
// listed slot; thread t takes row t of each, with every load of the R rows
// issued before any use (rows past the end load the last row and are
// dropped), so one slot lookup serves R rows and their loads overlap.  The
// slots of a work-group's logical blocks (at most kElemSlots: the launcher
// sizes the grid so) and the slot scalars the op uses are staged in LDS at
// the start, so a logical block waits on its row loads alone.
// Z: the inner vectors' type (float, or double for an FP64 inner cycle).
template <int OP, int R, typename Z>
__global__ __launch_bounds__(256) void k_b_outer_elem(BClassTab tab, BWork w, const SlotScalars *sc) {
    static_assert(OP == OP_SCALE || OP == OP_ACC || OP == OP_DOT_RZ || OP == OP_UPDATE,
                  "element-wise op");
    constexpr bool REDUCE = OP == OP_DOT_RZ || OP == OP_UPDATE;
    __shared__ int64_t pre[kMaxClasses + 1];
    if (threadIdx.x == 0) {
        int64_t acc = 0;
        pre[0] = 0;
        for (int c = 0; c < tab.nclass; ++c) {
            acc += (int64_t)(w.bounds[c + 1] - w.bounds[c]) * ((tab.c[c].bps + R - 1) / R);
            pre[c + 1] = acc;
        }
    }
    __syncthreads();
    const int64_t total = pre[tab.nclass];
    // per logical block: the slot, two scalars (UPDATE: alpha, 1/|r| of the
    // check; SCALE: 1/|r|; ACC, DOT_RZ: the cycle's scale) and a flag
    // (UPDATE: the current direction buffer; SCALE: r recomputed; ACC: the
    // first iteration)
    __shared__ int32_t tslot[kElemSlots], tflag[kElemSlots];
    __shared__ double ta[kElemSlots], tb[kElemSlots];
    const int64_t lb0 = first_block();
    const int64_t mine = lb0 < total ? (total - 1 - lb0) / gridDim.x + 1 : 0;
    for (int64_t k = threadIdx.x; k < mine; k += blockDim.x) {
        const int64_t lb = lb0 + k * gridDim.x;
        int c = 0;
        while (lb >= pre[c + 1]) ++c;
        const int64_t g = w.list[w.bounds[c] + (lb - pre[c]) / ((tab.c[c].bps + R - 1) / R)];
        const SlotScalars &st = sc[g];
        tslot[k] = (int32_t)g;
        ta[k] = OP == OP_UPDATE ? st.alpha : (OP == OP_SCALE ? st.inv : st.scale);
        tb[k] = st.inv_w;
        tflag[k] = OP == OP_UPDATE ? (st.pcur & 1) : (OP == OP_SCALE ? st.fresh : (st.iters == 1));
    }
    __syncthreads();
    int64_t kt = 0;
    for (int64_t lb = lb0; lb < total; lb += gridDim.x, ++kt) {
        int c = 0;
        while (lb >= pre[c + 1]) ++c;
        const BClassEntry &e = tab.c[c];
        int64_t si, sb;   // slot of the class list, group of R row blocks
        div32(lb - pre[c], (e.bps + R - 1) / R, si, sb);
        const int64_t g = tslot[kt];
        const int64_t ls = g - e.first;
        const int32_t flag = tflag[kt];
        // a running CG slot's rhs was written by OP_UPDATE (uniform per block)
        if (OP == OP_SCALE && !flag) continue;
        const double *rs = e.r + ls * e.vld, *qs = e.q + ls * e.vld;
        const double *ps = e.p + ls * e.vld + (OP == OP_UPDATE ? (int64_t)flag * e.palt : 0);
        const double *xs = e.x + ls * e.vld;
        const Z *zs = (const Z *)e.x_top + g * e.vld_top;
        const double alpha = ta[kt], inv = ta[kt], scale = ta[kt], inv_w = tb[kt];
        const bool first_acc = flag != 0;
        double *const xw = e.x + ls * e.vld, *const rw = e.r + ls * e.vld;
        Z *const tw = (Z *)e.rhs_top + g * e.vld_top;
        const int64_t n = e.n;
        int64_t row[R];
        bool in[R];
        double rv[R], zv[R], pv[R], qv[R], xo[R], part[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {   // every load first
            row[j] = (sb * R + j) * 256 + threadIdx.x;
            in[j] = row[j] < n;
            const int64_t rc = in[j] ? row[j] : n - 1;
            if constexpr (OP == OP_SCALE || OP == OP_DOT_RZ || OP == OP_UPDATE) rv[j] = rs[rc];
            if constexpr (OP == OP_ACC || OP == OP_DOT_RZ) zv[j] = (double)zs[rc];
            if constexpr (OP == OP_UPDATE) pv[j] = ps[rc];
            if constexpr (OP == OP_UPDATE) qv[j] = qs[rc];
            if constexpr (OP == OP_ACC || OP == OP_UPDATE) xo[j] = xs[rc];
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            part[j] = 0.0;
            const int64_t vi = row[j];
            if constexpr (OP == OP_SCALE) {
                if (in[j]) tw[vi] = (Z)(rv[j] * inv);
            } else if constexpr (OP == OP_ACC) {
                const double base = first_acc ? 0.0 : xo[j];
                if (in[j]) xw[vi] = base + zv[j] * scale;
            } else if constexpr (OP == OP_DOT_RZ) {
                // selects, not branches: a load used only inside a branch is
                // moved into it, which serialises the row groups
                const double t = rv[j] * (zv[j] * scale);
                part[j] = in[j] ? t : 0.0;
            } else if constexpr (OP == OP_UPDATE) {
                // and the next cycle's right-hand side, scaled by 1/|r| of
                // this iteration's check (the norm of r_new is not known yet)
                const double r = rv[j] - alpha * qv[j], xn = xo[j] + alpha * pv[j];
                part[j] = in[j] ? r * r : 0.0;
                if (in[j]) {
                    xw[vi] = xn;
                    rw[vi] = r;
                    tw[vi] = (Z)(r * inv_w);
                }
            }
        }
        if constexpr (REDUCE) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const double v = wave_sum(part[j]);
                const int64_t chunk = (sb * R + j) * (256 / kWave) + threadIdx.x / kWave;
                if (threadIdx.x % kWave == 0 && chunk < e.nps) e.partials[ls * e.pst + chunk] = v;
            }
        }
    }
}

// One 1024-thread block per listed slot (grid-stride): reduce the slot's
// partials and update its scalars.
//   FIN_CHECK_DC: MGHierarchy::solve_mixed's decision (initial norm, scale 1
//     on the first cycle, stop when norm < tol * initial, give up at
//     max_outer or on a non-finite norm, else scale = norm);
//   FIN_CHECK_CG: the same test on CG's recursive residual, confirmed by a
//     recomputed b - A x (VERIFY) before the slot retires;
//   FIN_BETA: rho = r.z, beta = rho / rho_prev (0 after a (re)start);
//   FIN_ALPHA: alpha = rho / p.Ap.
template <int OP>
__global__ __launch_bounds__(1024) void k_b_finish(BClassTab tab, BWork w, SlotScalars *sc, double tol,
                                                   int32_t max_outer) {
    const int64_t a0 = active_first(w), na = active_count(w);
    for (int64_t a = blockIdx.x; a < na; a += gridDim.x) {
        const int64_t g = w.list[a0 + a];
        int c = 0;
        while (c + 1 < tab.nclass && g >= tab.c[c + 1].first) ++c;
        const BClassEntry &e = tab.c[c];
        const int64_t ls = g - e.first;
        const int64_t np = OP == FIN_CHECK_DC ? e.bps : (OP == FIN_ALPHA ? e.ntiles : e.nps);
        const double h = sum_partials(e.partials + ls * e.pst, np, 1, 0);
        if (threadIdx.x == 0) {
            SlotScalars s = sc[g];
            if constexpr (OP == FIN_CHECK_DC || OP == FIN_CHECK_CG) {
                const double nrm = sqrt(h);
                s.norm = nrm;
                if (s.state == SLOT_NEW) {
                    s.initial = nrm;
                    s.iters = 0;
                    s.rel = 1.0;
                    s.rho = 0.0;
                    if (nrm == 0.0) {             // b - A x0 = 0: x0 is the solution
                        s.rel = 0.0;
                        s.state = SLOT_IDLE;
                        s.status = SLOT_OK;
                    } else if (!isfinite(nrm)) {
                        s.state = SLOT_IDLE;
                        s.status = SLOT_NOT_CONVERGED;
                    } else {
                        s.scale = OP == FIN_CHECK_DC ? 1.0 : nrm;
                        s.inv = OP == FIN_CHECK_DC ? 1.0 : 1.0 / nrm;
                        s.scale_w = nrm;
                        s.inv_w = 1.0 / nrm;
                        s.fresh = 1;
                        s.iters = 1;
                        s.state = SLOT_ACTIVE;
                    }
                } else {
                    s.rel = nrm / s.initial;
                    const bool met = nrm < tol * s.initial;
                    if (met && (OP == FIN_CHECK_DC || s.state == SLOT_VERIFY)) {
                        s.state = SLOT_IDLE;
                        s.status = SLOT_OK;
                    } else if (met) {
                        s.state = SLOT_VERIFY;
                    } else if (s.iters >= max_outer || !isfinite(nrm)) {
                        s.state = SLOT_IDLE;
                        s.status = SLOT_NOT_CONVERGED;
                    } else {
                        // DC, and CG after a recomputed residual: OP_SCALE
                        // writes rhs = r / |r|; a running CG slot's rhs was
                        // written by OP_UPDATE as r * inv_w, so its z takes
                        // that scale (|r| of the previous check)
                        const bool recomputed = OP == FIN_CHECK_DC || s.state == SLOT_VERIFY;
                        if (s.state == SLOT_VERIFY) {   // recursive residual drifted: restart
                            s.state = SLOT_ACTIVE;
                            s.rho = 0.0;
                        }
                        if (recomputed) {
                            s.scale = nrm;
                            s.inv = 1.0 / nrm;
                        } else {
                            s.scale = s.scale_w;
                            s.inv = s.inv_w;
                        }
                        s.scale_w = nrm;
                        s.inv_w = 1.0 / nrm;
                        s.fresh = recomputed ? 1 : 0;
                        s.iters += 1;
                    }
                }
            } else if constexpr (OP == FIN_BETA) {
                s.beta = (s.rho > 0.0 && h > 0.0 && isfinite(h)) ? h / s.rho : 0.0;
                s.rho = h;
            } else {
                const double ap = h;
                s.pcur ^= 1;   // the direction just formed (k_b_dir_spmv) is now current
                if (ap > 0.0 && isfinite(ap) && s.rho > 0.0) {
                    s.alpha = s.rho / ap;
                } else {                           // breakdown: give up on this slot
                    s.alpha = 0.0;
                    s.state = SLOT_IDLE;
                    s.status = SLOT_NOT_CONVERGED;
                }
            }
            sc[g] = s;
        }
        __syncthreads();   // sum_partials reuses shared memory in the next round
    }
}

// Level-wise elementwise launches: logical block = 256 consecutive entries of
// one active slot.
#define FOR_SLOT_ENTRIES(w, n, BODY)                                              \
    do {                                                                          \
        const int64_t a0_ = active_first(w);                                     \
        const int64_t bpe_ = ((n) + 255) / 256;                                   \
        const int64_t total_ = active_count(w) * bpe_;                            \
        for (int64_t lb_ = first_block(); lb_ < total_; lb_ += gridDim.x) {      \
            int64_t s_, b_;                                                       \
            div32(lb_, bpe_, s_, b_);                                             \
            const int64_t g = w.list[a0_ + s_];                                   \
            const int64_t i = b_ * 256 + threadIdx.x;                             \
            if (i < (n)) {                                                        \
                BODY;                                                             \
            }                                                                     \
        }                                                                         \
    } while (0)

// Row-wise transfers of the batch (every border Dirichlet, as the batched
// V-cycle uses them).  Logical unit = 64 consecutive nodes of one grid row of
// one listed slot, one wavefront per unit and four per work-group: the row,
// and with it the coarse (or fine) rows it reads, is uniform across the
// wavefront, so a node's position needs no division and its loads are
// consecutive across lanes.
struct RowUnit {
    int64_t g;    // slot
    int32_t row;  // grid row
    int32_t col;  // first column of the unit
};
// Unit u of the launch over `rows` rows of `cols` columns per slot; the
// arithmetic is wavefront-uniform (scalar registers).
__device__ inline RowUnit row_unit(const BWork &w, int64_t u, int32_t rows, int32_t cols) {
    const int32_t segs = (cols + kWave - 1) / kWave;
    const uint32_t per = (uint32_t)rows * (uint32_t)segs;
    const uint32_t uu = (uint32_t)u;
    const uint32_t e = uu / per, rem = uu - e * per;
    const uint32_t row = rem / (uint32_t)segs, seg = rem - row * (uint32_t)segs;
    RowUnit ru;
    ru.g = w.list[active_first(w) + (int64_t)e];
    ru.row = (int32_t)row;
    ru.col = (int32_t)seg * kWave;
    return ru;
}
// units of a launch over up to max_slots slots
inline int64_t row_units(int64_t max_slots, int64_t rows, int64_t cols) {
    return max_slots * rows * ((cols + kWave - 1) / kWave);
}

// Each wavefront takes kRowUnitsPerWave consecutive units per round and
// issues every load of them before the first store, so a round costs one
// memory round trip for several units (lanes past a row's end, and units past
// the launch's last, load from a valid node of the first slot and store
// nothing).
constexpr int kRowUnitsPerWave = 4;

// rhs_c[g] = R d_f[g] for active slots present on the fine level: coarse node
// (r, c) of an interior row/column gathers its nine fine values and adds the
// eight weighted terms to the centre value in the order restrict_node's
// sorting network puts them (the CPU sweep's order): the row terms A, G, H,
// B, E, F in that order, the column terms D and C placed by c - r (tags
// A 8r, G 8r+6, H 8r+7, B 8r+9, E 8r+12, F 8r+13, D 8c+3, C 8c+10); border
// nodes are Dirichlet (0).  The same bits as restrict_node with every border
// Dirichlet.
template <typename V>
__global__ __launch_bounds__(256) void k_b_restrict_rows(BWork w, V *rhs_c, int64_t vld_c, const V *d_f,
                                                         int64_t vld_f, int32_t m) {
    constexpr int U = kRowUnitsPerWave;
    const int32_t nc = m + 1, nf = 2 * m + 1;
    const int64_t units = active_count(w) * (int64_t)nc * ((nc + kWave - 1) / kWave);
    const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    for (int64_t lb = first_block(); lb * 4 * U < units; lb += gridDim.x) {
        const int64_t u0 = __builtin_amdgcn_readfirstlane((int32_t)((lb * 4 + wv) * U));
        V f[U][9];
        int64_t dst[U];
        int32_t dd[U];
        bool store[U], border[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const bool uok = u0 + k < units;
            const RowUnit ru = row_unit(w, uok ? u0 + k : 0, nc, nc);
            const int32_t c = ru.col + lane;
            store[k] = uok && c < nc;
            const int32_t r = ru.row;
            border[k] = r == 0 || r == m || c == 0 || c >= m;
            // interior position for the loads (a border node's are unused)
            const int32_t ir = r < 1 ? 1 : (r > m - 1 ? m - 1 : r);
            const int32_t ic = c < 1 ? 1 : (c > m - 1 ? m - 1 : c);
            dd[k] = c - r;
            dst[k] = ru.g * vld_c + (int64_t)r * nc + c;
            const V *fm = d_f + ru.g * vld_f + (int64_t)(2 * ir - 1) * nf;   // fine rows 2r - 1, 2r, 2r + 1
            const V *f0r = fm + nf, *fp = f0r + nf;
            const int32_t cc = 2 * ic;
            f[k][0] = fm[cc];       // A
            f[k][1] = fp[cc];       // B
            f[k][2] = f0r[cc + 1];  // C
            f[k][3] = f0r[cc - 1];  // D
            f[k][4] = fp[cc + 1];   // E
            f[k][5] = fp[cc - 1];   // F
            f[k][6] = fm[cc + 1];   // G
            f[k][7] = fm[cc - 1];   // H
            f[k][8] = f0r[cc];      // centre
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const V h = V(0.5), q = V(0.25);
            const V A = h * f[k][0], B = h * f[k][1], C = h * f[k][2], D = h * f[k][3];
            const V E = q * f[k][4], F = q * f[k][5], G = q * f[k][6], Hh = q * f[k][7];
            const int32_t d = dd[k];
            V s = f[k][8];
            s = d <= -1 ? s + D : s;
            s = d <= -2 ? s + C : s;
            s = s + A;
            s = d == -1 ? s + C : s;
            s = d == 0 ? s + D : s;
            s = s + G;
            s = s + Hh;
            s = s + B;
            s = d == 0 ? s + C : s;
            s = d == 1 ? s + D : s;
            s = s + E;
            s = s + F;
            s = d >= 2 ? s + D : s;
            s = d >= 1 ? s + C : s;
            if (store[k]) rhs_c[dst[k]] = border[k] ? V(0) : s;
        }
    }
}

// out_f[g] = x_f[g] + P x_c[g] (out_f may be x_f): fine node (R, C) from
// its up to four coarse neighbours with prolong_node's terms and order
// (border nodes Dirichlet: 0 correction).  The same bits as prolong_node
// with every border Dirichlet.
template <typename V>
__global__ __launch_bounds__(256) void k_b_prolong_rows(BWork w, V *x_f, int64_t vld_f, const V *x_c, int64_t vld_c,
                                                        int32_t m, V *out_f) {
    constexpr int U = kRowUnitsPerWave;
    const int32_t nc = m + 1, nf = 2 * m + 1;
    const int64_t units = active_count(w) * (int64_t)nf * ((nf + kWave - 1) / kWave);
    const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    for (int64_t lb = first_block(); lb * 4 * U < units; lb += gridDim.x) {
        const int64_t u0 = __builtin_amdgcn_readfirstlane((int32_t)((lb * 4 + wv) * U));
        V t[U][4], xv[U];
        int64_t id[U];
        int32_t Rk[U], Ck[U];
        bool store[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const bool uok = u0 + k < units;
            const RowUnit ru = row_unit(w, uok ? u0 + k : 0, nf, nf);
            const int32_t R = ru.row, C0 = ru.col + lane;
            store[k] = uok && C0 < nf;
            const int32_t C = C0 < nf ? C0 : nf - 1;
            Rk[k] = R;
            Ck[k] = C;
            prolong_taps(x_c + ru.g * vld_c, nc, R, C, t[k]);
            id[k] = ru.g * vld_f + (int64_t)R * nf + C;
            xv[k] = x_f[id[k]];
        }
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (store[k]) out_f[id[k]] = xv[k] + prolong_value(Rk[k], Ck[k], nf, t[k]) * V(1);
    }
}

template <typename V>
__global__ __launch_bounds__(256) void k_b_copy(BWork w, V *dst, const V *src, int64_t vld, int64_t n) {
    FOR_SLOT_ENTRIES(w, n, dst[g * vld + i] = src[g * vld + i]);
}

template <typename VT, typename V = typename VecOf<VT>::type>
__global__ __launch_bounds__(256) void k_b_coarse_cg(EllView A, BWork w, int64_t ax_ld, int64_t vld, const V *rhs,
                                                     V *x, V *g, V *u, V *en, double konv, int max_iters) {
    const int64_t a0 = active_first(w), na = active_count(w);
    const void *Ax0 = A.Ax;
    for (int64_t a = blockIdx.x; a < na; a += gridDim.x) {
        const int64_t s = w.list[a0 + a];
        A.Ax = (const VT *)Ax0 + s * ax_ld;
        const int64_t o = s * vld;
        (void)coarse_cg_block<VT>(A, rhs + o, x + o, g + o, u + o, en + o, konv, max_iters);
        __syncthreads();
    }
}

// ---- the coarse end of the V-cycle, one work-group per slot -------------------------
// One sweep of level L of slot `ls` by the whole work-group (T lanes per row;
// the loop bound is a multiple of blockDim so every lane of a row joins its
// lane reduction).  Same row sum and epilogue as k_b_ell.
template <typename VT, int T, int MODE, typename V = typename VecOf<VT>::type>
__device__ inline void fused_sweep(const FusedLevel &L, int64_t ls, V *out, const V *x, const V *rhs, const V *dinv) {
    const VT *Ax = (const VT *)L.Ax + ls * L.ax_ld;
    const int64_t vo = ls * L.vld;
    const V *xs = x + vo, *rs = rhs + vo, *ds = dinv + vo;
    V *os = out + vo;
    const int64_t total = ((L.n * T + blockDim.x - 1) / blockDim.x) * blockDim.x;
    for (int64_t gid = threadIdx.x; gid < total; gid += blockDim.x) {
        const int lane = (int)(gid & (T - 1));
        const int64_t row = gid / T;
        const bool active = row < L.n;
        V s = V(0);
        if (active) {
            if constexpr (MODE == ELL_JACOBI0)
                s = ell_lane_sum<VT, T, true, true>(L.Aj, Ax, row, lane, L.Arl[row], L.stride, rs, ds);
            else
                s = ell_lane_sum<VT, T, false, true>(L.Aj, Ax, row, lane, L.Arl[row], L.stride, xs);
        }
        s = lane_group_sum<T>(s);
        if (active && lane == 0) {
            if constexpr (MODE == ELL_DEFECT) {
                os[row] = rs[row] - s;
            } else {
                const V xin = (MODE == ELL_JACOBI0) ? ds[row] * rs[row] : xs[row];
                const V t = rs[row] - s;
                os[row] = xin + t * ds[row];
            }
        }
    }
}

template <typename VT, int MODE, typename V = typename VecOf<VT>::type>
__device__ inline void fused_sweep_t(const FusedLevel &L, int64_t ls, V *out, const V *x, const V *rhs,
                                     const V *dinv) {
    switch (L.T) {
        case 1: fused_sweep<VT, 1, MODE>(L, ls, out, x, rhs, dinv); break;
        case 2: fused_sweep<VT, 2, MODE>(L, ls, out, x, rhs, dinv); break;
        case 4: fused_sweep<VT, 4, MODE>(L, ls, out, x, rhs, dinv); break;
        default: fused_sweep<VT, 8, MODE>(L, ls, out, x, rhs, dinv); break;
    }
    __syncthreads();
}

// Level l of the cycle: fc's arrays, or (lds non-null) the LDS copies at
// the offsets fc.lo[l] (a copy of one level, so no runtime-indexed local array).
template <bool LDSM>
__device__ inline FusedLevel level_view(const FusedCycle &fc, int l, char *lds) {
    FusedLevel L = fc.lv[l];
    if constexpr (LDSM) {
        const FusedLds &o = fc.lo[l];
        L.Aj = (const int32_t *)(lds + o.aj);
        L.Arl = (const uint32_t *)(lds + o.arl);
        L.Ax = lds + o.ax;
        L.dinv = lds + o.dinv;
        L.rhs = lds + o.rhs;
        L.x = lds + o.x;
        L.d = lds + o.d;
        L.t = lds + o.t;
    }
    return L;
}

// One V-cycle of slot g from level `top` down to the coarse CG and back
// (fc's arrays indexed by slot g, or with LDSM the LDS copies at lds, g = 0;
// a compile-time choice, so the LDS form's accesses are LDS instructions).
template <typename VT, bool LDSM>
__device__ inline void fused_cycle_body(const FusedCycle &fc, int64_t g, int top, char *lds) {
    typedef typename VecOf<VT>::type V;
    const Mask8 mk{0xffu};   // all Dirichlet, as the level-by-level transfers
    // down: n_pre sweeps from zero (first two fused), defect, restriction
    for (int l = top; l >= 2; --l) {
        const FusedLevel L = level_view<LDSM>(fc, l, lds);
        V *Lx = (V *)L.x, *Lt = (V *)L.t, *Ld = (V *)L.d;
        const V *Lr = (const V *)L.rhs, *Ldi = (const V *)L.dinv;
        const int rest = fc.n_pre - 2;
        V *src = (rest % 2 == 0) ? Lx : Lt;
        fused_sweep_t<VT, ELL_JACOBI0>(L, g, src, (const V *)nullptr, Lr, Ldi);
        for (int k = 0; k < rest; ++k) {
            V *dst = (src == Lx) ? Lt : Lx;
            fused_sweep_t<VT, ELL_JACOBI>(L, g, dst, src, Lr, Ldi);
            src = dst;
        }
        fused_sweep_t<VT, ELL_DEFECT>(L, g, Ld, Lx, Lr, (const V *)nullptr);
        const FusedLevel C = level_view<LDSM>(fc, l - 1, lds);
        const int64_t m = C.side - 1;
        for (int64_t i = threadIdx.x; i < C.n; i += blockDim.x)
            restrict_node<V>((V *)C.rhs + g * C.vld, Ld + g * L.vld, m, mk, i);
        __syncthreads();
    }
    {   // coarse CG on level 1
        const FusedLevel L = level_view<LDSM>(fc, 1, lds);
        EllView A{};
        A.rows = L.n;
        A.cols = L.n;
        A.stride = L.stride;
        A.ncpr = L.ncpr;
        A.threads = L.T;
        A.Aj = L.Aj;
        A.Ax = (const VT *)L.Ax + g * L.ax_ld;
        A.Arl = L.Arl;
        const int64_t o = g * L.vld;
        V *cg = (V *)(LDSM ? lds + fc.og : fc.g), *cu = (V *)(LDSM ? lds + fc.ou : fc.u);
        V *cen = (V *)(LDSM ? lds + fc.oen : fc.en);
        (void)coarse_cg_block<VT>(A, (const V *)L.rhs + o, (V *)L.x + o, cg + o, cu + o, cen + o, fc.konv,
                                  fc.max_iters);
        __syncthreads();
    }
    // up: x += P x_coarse, n_post sweeps (result in x)
    for (int l = 2; l <= top; ++l) {
        const FusedLevel L = level_view<LDSM>(fc, l, lds);
        const FusedLevel C = level_view<LDSM>(fc, l - 1, lds);
        V *Lx = (V *)L.x, *Lt = (V *)L.t;
        const V *Lr = (const V *)L.rhs, *Ldi = (const V *)L.dinv;
        const int64_t m = C.side - 1;
        for (int64_t i = threadIdx.x; i < L.n; i += blockDim.x)
            prolong_node<V, true>(nullptr, (const V *)C.x + g * C.vld, m, mk, Lx + g * L.vld, V(1), i);
        __syncthreads();
        V *src = Lx, *dst = Lt;
        for (int k = 0; k < fc.n_post; ++k) {
            fused_sweep_t<VT, ELL_JACOBI>(L, g, dst, src, Lr, Ldi);
            V *tmp = src;
            src = dst;
            dst = tmp;
        }
        if (src != Lx) {
            for (int64_t i = threadIdx.x; i < L.n; i += blockDim.x) Lx[g * L.vld + i] = src[g * L.vld + i];
            __syncthreads();
        }
    }
}

template <typename VT>
__global__ __launch_bounds__(kFusedThreads) void k_b_coarse_vcycle(FusedCycle fc, BWork w) {
    const int64_t a0 = active_first(w), na = active_count(w);
    for (int64_t a = blockIdx.x; a < na; a += gridDim.x) {
        const int64_t g = w.list[a0 + a];
        int top = fc.top;
        while (top > 1 && g >= fc.lv[top].cap) --top;   // slots whose finest level is lower
        fused_cycle_body<VT, false>(fc, g, top, nullptr);
    }
}

// The same cycle with the slot's levels in LDS (dynamic LDS, offsets
// fc.lo, fc.og/ou/oen): the index arrays once per launch, then per slot its
// values, dinv and finest right-hand side in, the cycle on the LDS copies
// (slot index 0, the same code and arithmetic, so the same bits), and the
// finest level's x back out.  Every phase then waits on LDS instead of an L2
// round trip.
__device__ inline void lds_copy_words(void *dst, const void *src, int64_t bytes) {
    uint32_t *d = (uint32_t *)dst;
    const uint32_t *s = (const uint32_t *)src;
    for (int64_t i = threadIdx.x; i < bytes / 4; i += blockDim.x) d[i] = s[i];
}
template <typename VT>
__global__ __launch_bounds__(kFusedThreads) void k_b_coarse_vcycle_lds(FusedCycle fc, BWork w) {
    typedef typename VecOf<VT>::type V;
    extern __shared__ __align__(16) char lds[];
    const int64_t a0 = active_first(w), na = active_count(w);
    if ((int64_t)blockIdx.x >= na) return;   // uniform per work-group
    for (int l = 1; l <= fc.top; ++l) {
        const FusedLevel &L = fc.lv[l];
        lds_copy_words(lds + fc.lo[l].aj, L.Aj, 4 * L.ncpr * L.stride);
        lds_copy_words(lds + fc.lo[l].arl, L.Arl, 4 * L.n);
    }
    for (int64_t a = blockIdx.x; a < na; a += gridDim.x) {
        const int64_t g = w.list[a0 + a];
        int top = fc.top;
        while (top > 1 && g >= fc.lv[top].cap) --top;   // slots whose finest level is lower
        for (int l = 1; l <= top; ++l) {
            const FusedLevel &L = fc.lv[l];
            lds_copy_words(lds + fc.lo[l].ax, (const VT *)L.Ax + g * L.ax_ld, (int64_t)sizeof(VT) * L.ax_ld);
            lds_copy_words(lds + fc.lo[l].dinv, (const V *)L.dinv + g * L.vld, (int64_t)sizeof(V) * L.n);
        }
        lds_copy_words(lds + fc.lo[top].rhs, (const V *)fc.lv[top].rhs + g * fc.lv[top].vld,
                       (int64_t)sizeof(V) * fc.lv[top].n);
        __syncthreads();
        fused_cycle_body<VT, true>(fc, 0, top, lds);
        lds_copy_words((V *)fc.lv[top].x + g * fc.lv[top].vld, lds + fc.lo[top].x, (int64_t)sizeof(V) * fc.lv[top].n);
        __syncthreads();
    }
}

// Fresh scalars in `state` for the slots of one id list (passed by value).
__global__ void k_b_set_states(SlotScalars *sc, SlotIds ids, int32_t state) {
    const int k = (int)threadIdx.x;
    if (k < ids.n) {
        SlotScalars s = {};
        s.state = state;
        s.status = SLOT_PENDING;
        s.scale = 1.0;
        s.inv = 1.0;
        sc[ids.id[k]] = s;
    }
}

__global__ void k_b_fill_state(SlotScalars *sc, int64_t n, int32_t state) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g < n) {
        SlotScalars s = {};
        s.state = state;
        s.status = SLOT_PENDING;
        s.scale = 1.0;
        s.inv = 1.0;
        sc[g] = s;
    }
}

// ---- bulk hierarchy build: blockIdx.y = entry ------------------------------------
// finest coefficient and b (copied or drawn), start vector x0 (b on the
// first row, last row and first column, else 0)
__global__ void k_bb_fields(const BuildEntry *entries, BuildParams P) {
    const BuildEntry &e = entries[blockIdx.y];
    const int64_t side = P.side[e.level], n = side * side, cells = (side - 1) * (side - 1);
    double *kap = e.kappa[e.level];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (i < cells) kap[i] = e.seeded ? lognormal_at(P.sigma, e.coeff_seed, (uint64_t)i) : e.coeff_src[i];
        const double bi = e.seeded ? uniform_at(-1.0, 1.0, e.rhs_seed, (uint64_t)i) : e.b_src[i];
        e.b[i] = bi;
        e.x[i] = guess_node(i, n, side) ? bi : 0.0;
    }
}

// coefficient of level l from level l + 1, for entries finer than l
__global__ void k_bb_coarsen(const BuildEntry *entries, BuildParams P, int l) {
    const BuildEntry &e = entries[blockIdx.y];
    if (e.level <= l) return;
    const int64_t nc = P.side[l] - 1, cells = nc * nc;
    for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < cells; id += (int64_t)gridDim.x * blockDim.x)
        coarsen_cell(e.kappa[l + 1], 2 * nc, e.kappa[l], nc, P.averaging, id);
}

// inner values and dinv of level l (shared Aj / Arl untouched); a level
// whose values are not stored (e.Ax[l] null) gets its dinv only
template <typename VT>
__global__ void k_bb_level(const BuildEntry *entries, BuildParams P, int l) {
    const BuildEntry &e = entries[blockIdx.y];
    if (e.level < l) return;
    const int T = P.T[l];
    const int64_t rows = P.stride[l] / T, n = P.side[l] * P.side[l];
    typedef typename VecOf<VT>::type V;
    V *dinv = (V *)e.dinv[l];
    VT *st = (VT *)e.st[l];
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        if (e.Ax[l]) {
            q1_row<VT>(P.side[l], T, r, nullptr, (VT *)e.Ax[l], nullptr, P.stride[l], e.kappa[l], dinv, P.omega);
        } else if (r < n) {
            q1_dinv_row<VT>(P.side[l], r, e.kappa[l], dinv, P.omega);
            if (st) q1_stencil_node<VT>(P.side[l], r, e.kappa[l], st);
        }
    }
}

// FP64 outer values of the finest level (entries without stored FP64
// values are skipped)
__global__ void k_bb_outer(const BuildEntry *entries, BuildParams P) {
    const BuildEntry &e = entries[blockIdx.y];
    if (!e.Ax64) return;
    const int L = e.level;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < P.stride64[L];
         r += (int64_t)gridDim.x * blockDim.x)
        q1_row<double>(P.side[L], 1, r, nullptr, e.Ax64, nullptr, P.stride64[L], e.kappa[L], nullptr, 0.0);
}

// distinct-weight arrays of slot blockIdx.y from its cell coefficients
template <typename VT>
__global__ void k_stencil_fill(int64_t side, const double *kappa, int64_t kld, void *st, int64_t vld) {
    const int64_t g = blockIdx.y, n = side * side;
    VT *sg = (VT *)st + g * 4 * vld;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        q1_stencil_node<VT>(side, r, kappa + g * kld, sg);
}

typedef void (*BEllFn)(BMat, BWork, void *, const void *, const void *, const void *);

template <typename VT, int T, int G>
BEllFn pick_mode(EllMode mode) {
    switch (mode) {
        case ELL_DEFECT: return k_b_ell<VT, T, ELL_DEFECT, G>;
        case ELL_JACOBI: return k_b_ell<VT, T, ELL_JACOBI, G>;
        case ELL_JACOBI0: return k_b_ell<VT, T, ELL_JACOBI0, G>;
        default: return nullptr;
    }
}

template <typename VT, int T>
BEllFn pick_group(int group, EllMode mode) {
    switch (group) {
        case 1: return pick_mode<VT, T, 1>(mode);
        case 2: return pick_mode<VT, T, 2>(mode);
        case 4: return pick_mode<VT, T, 4>(mode);
        default: return nullptr;
    }
}

template <typename VT>
BEllFn pick_threads(int T, int group, EllMode mode) {
    switch (T) {
        case 1: return pick_group<VT, 1>(group, mode);
        case 2: return pick_group<VT, 2>(group, mode);
        case 4: return pick_group<VT, 4>(group, mode);
        case 8: return pick_group<VT, 8>(group, mode);
        default: return nullptr;
    }
}

int entry_grid(int64_t max_slots, int64_t n) { return batch_grid(max_slots * ((n + 255) / 256), 256); }
bool entry_fits(int64_t max_slots, int64_t n) { return max_slots * ((n + 255) / 256) <= kMaxLogicalBlocks; }

}  // namespace

int batch_grid(int64_t items, int block) {
    const int64_t cap = kResidentThreads / block;
    return (int)(items < 1 ? 1 : (items > cap ? cap : items));
}

hipError_t b_compact(const SlotScalars *sc, int64_t nslots, const int64_t *boundaries, int nb, int32_t *list,
                     int32_t *bounds, int mode, hipStream_t s) {
    if (nb < 1 || nb > kMaxClasses + 1 || nslots < 1) return hipErrorInvalidValue;
    BBoundaries bd{};
    bd.n = nb;
    for (int k = 0; k < nb; ++k) bd.v[k] = boundaries[k];
    hipLaunchKernelGGL(k_b_compact, dim3(1), dim3(1024), 0, s, sc, nslots, bd, list, bounds, mode);
    return hipGetLastError();
}

hipError_t b_ell(ValueType vt, int T, int block, int group, EllMode mode, const BMat &m, const BWork &w,
                 int64_t max_slots, void *out, const void *x, const void *rhs, const void *dinv, hipStream_t s) {
    BEllFn fn = nullptr;
    if (vt == VT_F32) fn = pick_threads<float>(T, group, mode);
    else if (vt == VT_BF16) fn = pick_threads<bf16>(T, group, mode);
    else if (vt == VT_F64) fn = pick_threads<double>(T, group, mode);
    if (!fn || block % kWave || block > 1024 || block < T) return hipErrorInvalidValue;
    if (m.bps != (m.n * T + block - 1) / block) return hipErrorInvalidValue;
    if (max_slots <= 0) return hipSuccess;
    const int64_t items = ((max_slots + group - 1) / group) * m.bps;
    if (items > kMaxLogicalBlocks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3(batch_grid(items, block)), dim3(block), 0, s, m, w, out, x, rhs, dinv);
    return hipGetLastError();
}

// Slots per logical block of the row-sum launches, 256-row blocks per
// logical block of the element-wise launches.
constexpr int kOuterRowGroup = 1, kOuterElemBlocks = 4;

hipError_t b_outer(OuterOp op, const BClassTab &tab, const BWork &w, int64_t max_blocks, const SlotScalars *sc,
                   hipStream_t s, bool mf) {
    if (max_blocks <= 0) return hipSuccess;
    if (max_blocks > kMaxLogicalBlocks) return hipErrorInvalidValue;
    if (mf)
        for (int c = 0; c < tab.nclass; ++c)
            if (!tab.c[c].kappa || tab.c[c].kld < (tab.c[c].side - 1) * (tab.c[c].side - 1)) return hipErrorInvalidValue;
    constexpr int G = kOuterRowGroup, R = kOuterElemBlocks;
    const int64_t ue = (max_blocks + R - 1) / R;
    const int ge_min = (int)((ue + kElemSlots - 1) / kElemSlots);   // at most kElemSlots logical blocks each
    const dim3 block(256), gr(batch_grid((max_blocks + G - 1) / G, 256)),
        ge(std::max(batch_grid(ue, 256), ge_min));
    switch (op) {
        case OP_DEFECT_ALL:
            if (mf)
                hipLaunchKernelGGL((k_b_outer_rows<OP_DEFECT_ALL, G, true, true>), gr, block, 0, s, tab, w, sc);
            else
                hipLaunchKernelGGL((k_b_outer_rows<OP_DEFECT_ALL, G, true, false>), gr, block, 0, s, tab, w, sc);
            break;
        case OP_DEFECT_SETTLE:
            if (mf)
                hipLaunchKernelGGL((k_b_outer_rows<OP_DEFECT_SETTLE, G, false, true>), gr, block, 0, s, tab, w, sc);
            else
                hipLaunchKernelGGL((k_b_outer_rows<OP_DEFECT_SETTLE, G, false, false>), gr, block, 0, s, tab, w, sc);
            break;
#define ELEM_LAUNCH(OPV)                                                                                 \
    do {                                                                                                 \
        if (tab.c[0].top64)                                                                              \
            hipLaunchKernelGGL((k_b_outer_elem<OPV, R, double>), ge, block, 0, s, tab, w, sc);          \
        else                                                                                             \
            hipLaunchKernelGGL((k_b_outer_elem<OPV, R, float>), ge, block, 0, s, tab, w, sc);           \
    } while (0)
        case OP_SCALE: ELEM_LAUNCH(OP_SCALE); break;
        case OP_ACC: ELEM_LAUNCH(OP_ACC); break;
        case OP_DOT_RZ: ELEM_LAUNCH(OP_DOT_RZ); break;
        case OP_UPDATE: ELEM_LAUNCH(OP_UPDATE); break;
#undef ELEM_LAUNCH
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t b_dir_spmv(const BClassTab &tab, const BWork &w, int64_t max_tiles, const SlotScalars *sc, hipStream_t s,
                      bool mf) {
    if (max_tiles <= 0) return hipSuccess;
    if (max_tiles > kMaxLogicalBlocks) return hipErrorInvalidValue;
    for (int c = 0; c < tab.nclass; ++c) {
        const BClassEntry &e = tab.c[c];
        const int64_t strips = (e.side + kPairWidth - 1) / kPairWidth;
        if (e.side < 3 || e.side > 65535 || e.strips != strips || e.tw != (e.side + strips - 1) / strips ||
            e.ntiles != strips * ((e.side + kDirRows - 1) / kDirRows) || e.ntiles > e.pst || !e.p || !e.q ||
            (mf && (!e.kappa || e.kld < (e.side - 1) * (e.side - 1))) || (!mf && !e.Ax))
            return hipErrorInvalidValue;
    }
    const int threads = kPairTileRows * kWave;
    const dim3 grid(tiled_grid(max_tiles, threads)), block(threads);
    if (tab.c[0].top64) {
        if (mf)
            hipLaunchKernelGGL((k_b_dir_spmv<true, double>), grid, block, 0, s, tab, w, sc);
        else
            hipLaunchKernelGGL((k_b_dir_spmv<false, double>), grid, block, 0, s, tab, w, sc);
    } else {
        if (mf)
            hipLaunchKernelGGL((k_b_dir_spmv<true, float>), grid, block, 0, s, tab, w, sc);
        else
            hipLaunchKernelGGL((k_b_dir_spmv<false, float>), grid, block, 0, s, tab, w, sc);
    }
    return hipGetLastError();
}

hipError_t b_finish(FinishOp op, const BClassTab &tab, const BWork &w, int64_t max_slots, SlotScalars *sc,
                    double tol, int32_t max_outer, hipStream_t s) {
    if (max_slots <= 0) return hipSuccess;
    const dim3 grid((unsigned)(max_slots < 1024 ? max_slots : 1024)), block(1024);
    switch (op) {
        case FIN_CHECK_DC:
            hipLaunchKernelGGL(k_b_finish<FIN_CHECK_DC>, grid, block, 0, s, tab, w, sc, tol, max_outer);
            break;
        case FIN_CHECK_CG:
            hipLaunchKernelGGL(k_b_finish<FIN_CHECK_CG>, grid, block, 0, s, tab, w, sc, tol, max_outer);
            break;
        case FIN_BETA: hipLaunchKernelGGL(k_b_finish<FIN_BETA>, grid, block, 0, s, tab, w, sc, tol, max_outer); break;
        case FIN_ALPHA:
            hipLaunchKernelGGL(k_b_finish<FIN_ALPHA>, grid, block, 0, s, tab, w, sc, tol, max_outer);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t b_restrict(ValueType vt, const BLevel &fine, const BLevel &coarse, const BWork &w, hipStream_t s) {
    const int64_t m = coarse.side - 1;
    if (m < 2 || fine.side != 2 * m + 1 || fine.cap > coarse.cap || fine.side > kMaxTransferSide)
        return hipErrorInvalidValue;
    const int64_t units = row_units(fine.cap, coarse.side, coarse.side);
    if (units > 0x7fffffff - 4 * kRowUnitsPerWave) return hipErrorInvalidValue;
    const dim3 grid(batch_grid((units + 4 * kRowUnitsPerWave - 1) / (4 * kRowUnitsPerWave), 256)), block(256);
    if (vt == VT_F64)
        hipLaunchKernelGGL(k_b_restrict_rows<double>, grid, block, 0, s, w, (double *)coarse.rhs, coarse.vld,
                           (const double *)fine.d, fine.vld, (int32_t)m);
    else
        hipLaunchKernelGGL(k_b_restrict_rows<float>, grid, block, 0, s, w, (float *)coarse.rhs, coarse.vld,
                           (const float *)fine.d, fine.vld, (int32_t)m);
    return hipGetLastError();
}

hipError_t b_prolong_axpy(ValueType vt, const BLevel &fine, const BLevel &coarse, const void *xc, const BWork &w,
                          void *out, hipStream_t s) {
    const int64_t m = coarse.side - 1;
    if (fine.side != 2 * m + 1 || fine.cap > coarse.cap || fine.side > kMaxTransferSide) return hipErrorInvalidValue;
    if ((out != fine.x && out != fine.t) || (xc != coarse.x && xc != coarse.t)) return hipErrorInvalidValue;
    const int64_t units = row_units(fine.cap, fine.side, fine.side);
    if (units > 0x7fffffff - 4 * kRowUnitsPerWave) return hipErrorInvalidValue;
    const dim3 grid(batch_grid((units + 4 * kRowUnitsPerWave - 1) / (4 * kRowUnitsPerWave), 256)), block(256);
    if (vt == VT_F64)
        hipLaunchKernelGGL(k_b_prolong_rows<double>, grid, block, 0, s, w, (double *)fine.x, fine.vld,
                           (const double *)xc, coarse.vld, (int32_t)m, (double *)out);
    else
        hipLaunchKernelGGL(k_b_prolong_rows<float>, grid, block, 0, s, w, (float *)fine.x, fine.vld,
                           (const float *)xc, coarse.vld, (int32_t)m, (float *)out);
    return hipGetLastError();
}

template <typename VT, int P, int WM, bool FP>
void pair_launch_rows(dim3 grid, dim3 block, hipStream_t s, const BMat &m, const BWork &w, const PairGeom &pg,
                      void *out0, void *out1, const void *x, const void *rhs, const void *dinv, const void *xc,
                      int64_t vld_c, int nr) {
    typedef typename VecOf<VT>::type V;
    if constexpr (sizeof(V) == 4) {   // FP64 vectors: two rows per thread would spill
        if (nr == 2) {
            hipLaunchKernelGGL((k_b_pair<VT, P, WM, FP, 2>), grid, block, 0, s, m, w, pg, (V *)out0, (V *)out1,
                               (const V *)x, (const V *)rhs, (const V *)dinv, (const V *)xc, vld_c);
            return;
        }
    }
    hipLaunchKernelGGL((k_b_pair<VT, P, WM, FP, 1>), grid, block, 0, s, m, w, pg, (V *)out0, (V *)out1,
                       (const V *)x, (const V *)rhs, (const V *)dinv, (const V *)xc, vld_c);
}

template <typename VT, int P, int WM>
void pair_launch(dim3 grid, dim3 block, hipStream_t s, const BMat &m, const BWork &w, const PairGeom &pg, void *out0,
                 void *out1, const void *x, const void *rhs, const void *dinv, const void *xc, int64_t vld_c, int nr) {
    if constexpr (P == PAIR_POST) {
        if (xc) {
            pair_launch_rows<VT, P, WM, true>(grid, block, s, m, w, pg, out0, out1, x, rhs, dinv, xc, vld_c, nr);
            return;
        }
    }
    pair_launch_rows<VT, P, WM, false>(grid, block, s, m, w, pg, out0, out1, x, rhs, dinv, nullptr, 0, nr);
}

hipError_t b_pair(ValueType vt, int pair, const BLevel &lv, const BWork &w, void *out0, void *out1,
                  const void *xin, hipStream_t s, int wmode, const BLevel *coarse, const void *xc) {
    if (lv.T != 1 || lv.side < 3 || lv.ncpr != 9 || lv.side > 65535) return hipErrorInvalidValue;
    if (pair != PAIR_PRE && pair != PAIR_POST) return hipErrorInvalidValue;
    if (!out0 || (pair == PAIR_PRE && !out1) || (pair == PAIR_POST && (!xin || xin == out0)))
        return hipErrorInvalidValue;
    const bool fp = xc != nullptr;   // the prolongation fused in
    if (fp && (pair != PAIR_POST || !coarse || lv.side != 2 * (coarse->side - 1) + 1 || lv.cap > coarse->cap ||
               lv.side > kMaxTransferSide || xc == out0))
        return hipErrorInvalidValue;
    const int64_t vld_c = fp ? coarse->vld : 0;
    if (wmode < 0 || wmode > 2) return hipErrorInvalidValue;
    if (wmode == 1 && (!lv.kappa || lv.kld < (lv.side - 1) * (lv.side - 1))) return hipErrorInvalidValue;
    if (wmode == 2 && !lv.st) return hipErrorInvalidValue;
    if (wmode == 0 && (!lv.Ax || !lv.Aj)) return hipErrorInvalidValue;
    PairGeom pg;
    pg.side = (int32_t)lv.side;
    pg.strips = (int32_t)((lv.side + kPairWidth - 1) / kPairWidth);
    pg.width = (int32_t)((lv.side + pg.strips - 1) / pg.strips);
    // two grid rows per thread on the larger levels (taller tiles: half the
    // halo rows per node, twice the loads in flight per thread)
    const int nr = lv.side >= kPairTallMinSide && vt != VT_F64 ? 2 : 1;
    const int64_t rows = (int64_t)kPairTileRows * nr - 2;
    pg.segs = (int32_t)((lv.side + rows - 1) / rows);
    if (lv.cap <= 0) return hipSuccess;
    const int64_t units = lv.cap * (int64_t)pg.strips * pg.segs;
    if (units > kMaxLogicalBlocks) return hipErrorInvalidValue;
    const BMat m{lv.Aj, lv.Arl, lv.Ax, lv.stride, lv.ax_ld, lv.n, lv.vld, 0, lv.kappa, lv.kld, lv.st};
    const int threads = kPairTileRows * kWave;
    const dim3 grid(tiled_grid(units, threads)), block(threads);
    const void *x = pair == PAIR_POST ? xin : lv.rhs;   // unused by PAIR_PRE
#define PAIR_LAUNCH(VT, P, MFV) \
    pair_launch<VT, P, MFV>(grid, block, s, m, w, pg, out0, out1, x, lv.rhs, lv.dinv, xc, vld_c, nr)
#define PAIR_WM(VT, P)                                                        \
    do {                                                                      \
        if (wmode == 2)                                                       \
            PAIR_LAUNCH(VT, P, 2);                                            \
        else if (wmode == 1)                                                  \
            PAIR_LAUNCH(VT, P, 1);                                            \
        else                                                                  \
            PAIR_LAUNCH(VT, P, 0);                                            \
    } while (0)
#define PAIR_VT(VT)                                                           \
    do {                                                                      \
        if (pair == PAIR_PRE)                                                 \
            PAIR_WM(VT, PAIR_PRE);                                            \
        else                                                                  \
            PAIR_WM(VT, PAIR_POST);                                           \
    } while (0)
    if (vt == VT_F32)
        PAIR_VT(float);
    else if (vt == VT_BF16)
        PAIR_VT(bf16);
    else if (vt == VT_F64)
        PAIR_VT(double);
    else
        return hipErrorInvalidValue;
#undef PAIR_VT
#undef PAIR_WM
#undef PAIR_LAUNCH
    return hipGetLastError();
}

hipError_t b_copy(ValueType vt, const BLevel &lv, const BWork &w, void *dst, const void *src, hipStream_t s) {
    if (!entry_fits(lv.cap, lv.n)) return hipErrorInvalidValue;
    const dim3 grid(entry_grid(lv.cap, lv.n)), block(256);
    if (vt == VT_F64)
        hipLaunchKernelGGL(k_b_copy<double>, grid, block, 0, s, w, (double *)dst, (const double *)src, lv.vld, lv.n);
    else
        hipLaunchKernelGGL(k_b_copy<float>, grid, block, 0, s, w, (float *)dst, (const float *)src, lv.vld, lv.n);
    return hipGetLastError();
}

hipError_t b_coarse_cg(ValueType vt, const BLevel &lv, const BWork &w, hipStream_t s) {
    EllView A{};
    A.rows = lv.n;
    A.cols = lv.n;
    A.stride = lv.stride;
    A.ncpr = lv.ncpr;
    A.threads = lv.T;
    A.Aj = lv.Aj;
    A.Ax = lv.Ax;
    A.Arl = lv.Arl;
    A.vtype = vt;
    // the inner type's epsilon, as the single-system coarse solve
    const double konv = vt == VT_F64 ? 2.220446049250313e-16 : 1.1920928955078125e-07;
    const int max_iters = (int)(4 * lv.n + 64);
    const int grid = (int)(lv.cap < 2048 ? lv.cap : 2048);
    float *fr = (float *)lv.rhs, *fx = (float *)lv.x, *fg = (float *)lv.g, *fu = (float *)lv.u, *fe = (float *)lv.en;
    double *dr = (double *)lv.rhs, *dx = (double *)lv.x, *dg = (double *)lv.g, *du = (double *)lv.u,
           *de = (double *)lv.en;
    if (vt == VT_F32)
        hipLaunchKernelGGL(k_b_coarse_cg<float>, dim3(grid), dim3(256), 0, s, A, w, lv.ax_ld, lv.vld,
                           (const float *)fr, fx, fg, fu, fe, konv, max_iters);
    else if (vt == VT_BF16)
        hipLaunchKernelGGL(k_b_coarse_cg<bf16>, dim3(grid), dim3(256), 0, s, A, w, lv.ax_ld, lv.vld,
                           (const float *)fr, fx, fg, fu, fe, konv, max_iters);
    else if (vt == VT_F64)
        hipLaunchKernelGGL(k_b_coarse_cg<double>, dim3(grid), dim3(256), 0, s, A, w, lv.ax_ld, lv.vld,
                           (const double *)dr, dx, dg, du, de, konv, max_iters);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t b_coarse_vcycle(ValueType vt, const FusedCycle &fc, const BWork &w, int64_t max_slots, hipStream_t s) {
    if (max_slots <= 0) return hipSuccess;
    if (fc.top < 1 || fc.top > kMaxFusedLevels || fc.n_pre < 2 || fc.n_post < 0) return hipErrorInvalidValue;
    for (int l = 1; l <= fc.top; ++l) {
        const int T = fc.lv[l].T;
        if (T != 1 && T != 2 && T != 4 && T != 8) return hipErrorInvalidValue;
        if (l > 1 && (fc.lv[l].side != 2 * (fc.lv[l - 1].side - 1) + 1 || fc.lv[l].cap > fc.lv[l - 1].cap))
            return hipErrorInvalidValue;
    }
    const int grid = (int)(max_slots < 2048 ? max_slots : 2048);
    // the LDS form when every level of the cycle fits in 60 KiB (with the
    // kernel's static LDS, within a work-group's 64 KiB)
    FusedCycle f = fc;
    const size_t vs = vt_size(vt), es = vt == VT_F64 ? 8 : 4;
    int64_t at = 0;
    auto take = [&](int64_t bytes) {
        const int64_t o = at;
        at += (bytes + 15) / 16 * 16;
        return (int32_t)o;
    };
    for (int l = 1; l <= f.top; ++l) {
        const FusedLevel &L = f.lv[l];
        FusedLds &o = f.lo[l];
        o.aj = take(4 * L.ncpr * L.stride);
        o.arl = take(4 * L.n);
        o.ax = take((int64_t)vs * L.ax_ld);
        o.dinv = take((int64_t)es * L.n);
        o.rhs = take((int64_t)es * L.n);
        o.x = take((int64_t)es * L.n);
        o.d = take((int64_t)es * L.n);
        o.t = take((int64_t)es * L.n);
    }
    f.og = take((int64_t)es * f.lv[1].n);
    f.ou = take((int64_t)es * f.lv[1].n);
    f.oen = take((int64_t)es * f.lv[1].n);
    const bool lds = at <= 60 * 1024;
#define CYCLE_LAUNCH(VT)                                                                                  \
    do {                                                                                                  \
        if (lds)                                                                                          \
            hipLaunchKernelGGL(k_b_coarse_vcycle_lds<VT>, dim3(grid), dim3(kFusedThreads), (size_t)at, s, f, w); \
        else                                                                                              \
            hipLaunchKernelGGL(k_b_coarse_vcycle<VT>, dim3(grid), dim3(kFusedThreads), 0, s, f, w);      \
    } while (0)
    if (vt == VT_F32)
        CYCLE_LAUNCH(float);
    else if (vt == VT_BF16)
        CYCLE_LAUNCH(bf16);
    else if (vt == VT_F64)
        CYCLE_LAUNCH(double);
    else
        return hipErrorInvalidValue;
#undef CYCLE_LAUNCH
    return hipGetLastError();
}

hipError_t b_build(const BuildEntry *entries, const BuildParams &P, int max_level, hipStream_t s) {
    if (P.entries <= 0) return hipSuccess;
    if (P.entries > 65535 || max_level < 1 || max_level > kMaxLevel) return hipErrorInvalidValue;
    // blocks per entry: enough for the largest entry's work, but about 8192
    // blocks in all (each thread strides over more of an entry when many are
    // built at once), so the entries smaller than the largest do not launch
    // thousands of empty blocks each
    const int64_t per_cap = std::max<int64_t>(16, 8192 / P.entries);
    auto grid = [&](int64_t work) {
        int64_t g = (work + 255) / 256;
        g = std::min<int64_t>(g, std::min<int64_t>(2048, per_cap));
        return dim3((unsigned)(g < 1 ? 1 : g), (unsigned)P.entries);
    };
    const int64_t nmax = P.side[max_level] * P.side[max_level];
    hipLaunchKernelGGL(k_bb_fields, grid(nmax), dim3(256), 0, s, entries, P);
    for (int l = max_level - 1; l >= 1; --l) {
        const int64_t nc = P.side[l] - 1;
        hipLaunchKernelGGL(k_bb_coarsen, grid(nc * nc), dim3(256), 0, s, entries, P, l);
    }
    for (int l = max_level; l >= 1; --l) {
        const int64_t rows = P.stride[l] / P.T[l];
        if (P.vt == VT_F32)
            hipLaunchKernelGGL(k_bb_level<float>, grid(rows), dim3(256), 0, s, entries, P, l);
        else if (P.vt == VT_BF16)
            hipLaunchKernelGGL(k_bb_level<bf16>, grid(rows), dim3(256), 0, s, entries, P, l);
        else if (P.vt == VT_F64)
            hipLaunchKernelGGL(k_bb_level<double>, grid(rows), dim3(256), 0, s, entries, P, l);
        else
            return hipErrorInvalidValue;
    }
    if (P.outer_values) hipLaunchKernelGGL(k_bb_outer, grid(P.stride64[max_level]), dim3(256), 0, s, entries, P);
    return hipGetLastError();
}

hipError_t b_stencil_fill(ValueType vt, int64_t side, const double *kappa, int64_t kld, void *st, int64_t vld,
                          int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (side < 3 || count > 65535 || vld < side * side) return hipErrorInvalidValue;
    const int64_t n = side * side;
    const dim3 grid((unsigned)std::min<int64_t>(2048, (n + 255) / 256), (unsigned)count), block(256);
    switch (vt) {
        case VT_F32: hipLaunchKernelGGL(k_stencil_fill<float>, grid, block, 0, s, side, kappa, kld, st, vld); break;
        case VT_BF16: hipLaunchKernelGGL(k_stencil_fill<bf16>, grid, block, 0, s, side, kappa, kld, st, vld); break;
        case VT_F64: hipLaunchKernelGGL(k_stencil_fill<double>, grid, block, 0, s, side, kappa, kld, st, vld); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t b_set_states(SlotScalars *sc, const int64_t *slots, int64_t n, int32_t state, hipStream_t s) {
    for (int64_t k0 = 0; k0 < n; k0 += kMaxSlotIds) {
        SlotIds ids{};
        ids.n = (int32_t)(n - k0 < kMaxSlotIds ? n - k0 : kMaxSlotIds);
        for (int k = 0; k < ids.n; ++k) ids.id[k] = (int32_t)slots[k0 + k];
        hipLaunchKernelGGL(k_b_set_states, dim3(1), dim3(kMaxSlotIds), 0, s, sc, ids, state);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t b_fill_state(SlotScalars *sc, int64_t nslots, int32_t state, hipStream_t s) {
    if (nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_b_fill_state, dim3(ceil_div(nslots, 256)), dim3(256), 0, s, sc, nslots, state);
    return hipGetLastError();
}

}  // namespace hipx
}  // namespace honei
