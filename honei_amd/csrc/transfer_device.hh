// transfer_device.hh -- per-node geometric multigrid transfers on
// lexicographic (2^L+1)^2 Q1 grids, shared by the single-system and the
// batched transfer kernels.
//
// Restriction is the transpose of bilinear interpolation (weights 1, 1/2,
// 1/4; math/restriction.hh:40-172); prolongation is bilinear interpolation
// (math/prolongation.hh:41-139).  Both finish with the Dirichlet mask of
// math/apply_dirichlet_boundaries.hh:28-85 (8 flags: 4 corners, 4 edges;
// 2 = Dirichlet -> node set to 0).
//
// The reference accumulates each coarse node from 9 fine values in the order
// its sweep loops visit them.  restrict_node forms the same 9 terms, tags
// each with the (sweep, sub-step) at which the reference adds it, sorts the
// (at most 8) added terms by that tag with a small sorting network and sums
// in that order, so the result is bit-identical to the CPU path rather than
// merely close.
#pragma once

// Multiply and add stay separate roundings in everything that follows (the
// row sums and transfers reproduce the CPU path bit for bit).
#pragma clang fp contract(off)

#include "hip_common.hh"

namespace honei {
namespace hipx {

// Bit e set: border part e is Dirichlet (its nodes are set to 0).
struct Mask8 {
    uint32_t bits;
};

// Host: the reference's macroBorderMask (8 values, 2 = Dirichlet) packed for
// kernel arguments; a null mask means all Dirichlet.
inline Mask8 pack_mask(const uint64_t *mask) {
    Mask8 m{0u};
    for (int i = 0; i < 8; ++i)
        if (!mask || mask[i] == 2) m.bits |= 1u << i;
    return m;
}

// Grids handled by the transfers: node ids fit 32 bits (side <= 65535).
constexpr int64_t kMaxTransferSide = 65535;

__device__ inline bool dirichlet_zero(uint32_t r, uint32_t c, uint32_t M, const Mask8 &mk) {
    int e;
    if (r == 0 && c == 0) e = 0;
    else if (r == 0 && c == M - 1) e = 1;
    else if (r == M - 1 && c == M - 1) e = 2;
    else if (r == M - 1 && c == 0) e = 3;
    else if (r == 0) e = 4;
    else if (c == M - 1) e = 5;
    else if (r == M - 1) e = 6;
    else if (c == 0) e = 7;
    else return false;
    return (mk.bits >> e) & 1u;
}

template <typename V>
__device__ inline void cswap(int32_t &ka, V &va, int32_t &kb, V &vb) {
    if (kb < ka) {
        int32_t tk = ka; ka = kb; kb = tk;
        V tv = va; va = vb; vb = tv;
    }
}

// coarse node id (< (m+1)^2) of the restriction of `fine`.
template <typename V>
__device__ inline void restrict_node(V *__restrict__ coarse, const V *__restrict__ fine, int64_t m, const Mask8 &mk,
                                     int64_t id) {
    const int64_t nc = m + 1, nf = 2 * m + 1;
    const int64_t r = (uint32_t)id / (uint32_t)nc, c = (uint32_t)id % (uint32_t)nc;
    const bool dirichlet = dirichlet_zero((uint32_t)r, (uint32_t)c, (uint32_t)nc, mk);
    auto F = [&](int64_t fr, int64_t fc) { return fine[fr * nf + fc]; };
    const int32_t NONE = INT32_MAX;
    const int32_t ri = (int32_t)r, ci = (int32_t)c;
    // all nine fine values are loaded (rows / columns outside the grid
    // clamped to a valid neighbour) and the absent terms selected away, so the
    // loads are issued together
    const bool up = r >= 1, dn = r + 1 <= m, lf = c >= 1, rt = c + 1 <= m;
    const int64_t rm = up ? 2 * r - 1 : 2 * r, rp = dn ? 2 * r + 1 : 2 * r;
    const int64_t cm = lf ? 2 * c - 1 : 2 * c, cp = rt ? 2 * c + 1 : 2 * c;
    const V fA = F(rm, 2 * c), fB = F(rp, 2 * c), fC = F(2 * r, cp), fD = F(2 * r, cm);
    const V fE = F(rp, cp), fF = F(rp, cm), fG = F(rm, cp), fH = F(rm, cm), f0 = F(2 * r, 2 * c);
    // tag = sweep index i (1..m) * 8 + sub-step (A..H = 0..7)
    int32_t k[8];
    V v[8];
    const V h = V(0.5), q = V(0.25);
    // A_i (i=r): row i += .5 F(2i-1, 2c)
    k[0] = up ? ri * 8 + 0 : NONE;             v[0] = up ? h * fA : V(0);
    // B_i (i=r+1): row i-1 += .5 F(2i-1, 2c)
    k[1] = dn ? (ri + 1) * 8 + 1 : NONE;       v[1] = dn ? h * fB : V(0);
    // C_i (i=c+1): col i-1 += .5 F(2r, 2i-1)
    k[2] = rt ? (ci + 1) * 8 + 2 : NONE;       v[2] = rt ? h * fC : V(0);
    // D_i (i=c): col i += .5 F(2r, 2i-1)
    k[3] = lf ? ci * 8 + 3 : NONE;             v[3] = lf ? h * fD : V(0);
    // E_i (i=r+1, j=c<m): (i-1, j) += .25 F(2i-1, 2j+1)
    k[4] = (dn && c < m) ? (ri + 1) * 8 + 4 : NONE;
    v[4] = (dn && c < m) ? q * fE : V(0);
    // F_i (i=r+1, j=c-1): (i-1, j+1) += .25 F(2i-1, 2j+1)
    k[5] = (dn && lf) ? (ri + 1) * 8 + 5 : NONE;
    v[5] = (dn && lf) ? q * fF : V(0);
    // G_i (i=r, j=c<m): (i, j) += .25 F(2i-1, 2j+1)
    k[6] = (up && c < m) ? ri * 8 + 6 : NONE;
    v[6] = (up && c < m) ? q * fG : V(0);
    // H_i (i=r, j=c-1): (i, j+1) += .25 F(2i-1, 2j+1)
    k[7] = (up && lf) ? ri * 8 + 7 : NONE;
    v[7] = (up && lf) ? q * fH : V(0);
    // Batcher odd-even merge sort network for 8 keys (19 comparators)
    cswap(k[0], v[0], k[1], v[1]); cswap(k[2], v[2], k[3], v[3]);
    cswap(k[4], v[4], k[5], v[5]); cswap(k[6], v[6], k[7], v[7]);
    cswap(k[0], v[0], k[2], v[2]); cswap(k[1], v[1], k[3], v[3]);
    cswap(k[4], v[4], k[6], v[6]); cswap(k[5], v[5], k[7], v[7]);
    cswap(k[1], v[1], k[2], v[2]); cswap(k[5], v[5], k[6], v[6]);
    cswap(k[0], v[0], k[4], v[4]); cswap(k[1], v[1], k[5], v[5]);
    cswap(k[2], v[2], k[6], v[6]); cswap(k[3], v[3], k[7], v[7]);
    cswap(k[2], v[2], k[4], v[4]); cswap(k[3], v[3], k[5], v[5]);
    cswap(k[1], v[1], k[2], v[2]); cswap(k[3], v[3], k[4], v[4]);
    cswap(k[5], v[5], k[6], v[6]);
    V s = f0;
#pragma unroll
    for (int t = 0; t < 8; ++t) s = (k[t] != NONE) ? s + v[t] : s;
    coarse[id] = dirichlet ? V(0) : s;
}

// fine[R,C] from coarse, in the reference's accumulation order; optionally
// x[R,C] + alpha * fine[R,C] (the MG correction ScaledSum, fused) stored to
// xo[R,C] (xo null: back to x).
template <typename V, bool AXPY>
__device__ inline void prolong_node(V *__restrict__ fine, const V *__restrict__ coarse, int64_t m, const Mask8 &mk,
                                    V *x, V alpha, int64_t id, V *xo = nullptr) {
    const int64_t nc = m + 1, nf = 2 * m + 1;
    const int64_t R = (uint32_t)id / (uint32_t)nf, C = (uint32_t)id % (uint32_t)nf;
    // The up to four coarse neighbours a (rlo, clo), b (rlo, chi), c (rhi,
    // clo), d (rhi, chi) -- one coarse row / column when R / C is even -- are
    // all loaded, every case is formed and one is selected, so neighbouring
    // lanes (alternating parities) take no separate paths.  Each case keeps
    // the reference's terms and order: even-even a; even R a/2 + b/2; even C
    // a/2 + c/2; odd-odd a/4 + b/4 + c/4 + d/4 (each from a zero start).
    const bool oddR = (R & 1) != 0, oddC = (C & 1) != 0;
    const int64_t rlo = (R - (oddR ? 1 : 0)) / 2, rhi = (R + (oddR ? 1 : 0)) / 2;
    const int64_t clo = (C - (oddC ? 1 : 0)) / 2, chi = (C + (oddC ? 1 : 0)) / 2;
    const V a = coarse[rlo * nc + clo], b = coarse[rlo * nc + chi];
    const V c = coarse[rhi * nc + clo], d = coarse[rhi * nc + chi];
    const V h = V(0.5), q = V(0.25);
    const V f_rc = ((V(0) + a * q + b * q) + c * q) + d * q;
    const V f_c = (V(0) + a * h) + b * h;
    const V f_r = (V(0) + a * h) + c * h;
    V f = oddR ? (oddC ? f_rc : f_r) : (oddC ? f_c : a);
    if (dirichlet_zero((uint32_t)R, (uint32_t)C, (uint32_t)nf, mk)) f = V(0);
    if (fine) fine[id] = f;
    if constexpr (AXPY) (xo ? xo : x)[id] = x[id] + f * alpha;
}

}  // namespace hipx
}  // namespace honei
