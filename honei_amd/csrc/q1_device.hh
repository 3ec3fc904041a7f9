// q1_device.hh -- per-row / per-cell device code of the Q1 hierarchy build,
// shared by the single-system kernels (one system per launch) and the bulk
// build of the batched solver (many systems per launch), so both produce the
// same bits.
#pragma once

// Multiply and add stay separate roundings in everything that follows.
#pragma clang fp contract(off)

#include "exact_div.hh"
#include "hip_common.hh"

namespace honei {
namespace hipx {

template <typename VT, typename V>
__device__ inline void q1_put(VT *Ax, int64_t s, V v) { Ax[s] = (VT)v; }
template <>
__device__ inline void q1_put<bf16, float>(bf16 *Ax, int64_t s, float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    Ax[s].bits = (uint16_t)(u >> 16);
}

// The nine weights of interior node (i, j) of the variable-coefficient Q1
// operator (columns (i-1, j-1) .. (i+1, j+1) in column order): the bilinear
// element stiffness kappa/6 * [[4,-1,-2,-1],[-1,4,-1,-2],[-2,-1,4,-1],
// [-1,-2,-1,4]] summed over the four cells around the node, cells row-major
// in coeff[(side-1)^2], each weight formed in double and rounded once to V.
// q1_row stores these values; the matrix-free batched kernels form them
// again from the same coefficients, so both see the same bits.  Here from
// the four cell values around the node (sw = cell (i-1, j-1), se = (i-1, j),
// nw = (i, j-1), ne = (i, j)).
template <typename V>
__device__ inline void q1_cell_weights(double sw, double se, double nw, double ne, V (&w)[9]) {
    // the values of -c / 3.0, -(a + b) / 6.0 and 2.0 * (sum) / 3.0
    w[0] = (V)(-div3_rn(sw));                  // (i-1, j-1)
    w[1] = (V)(-0.5 * div3_rn(sw + se));       // (i-1, j)
    w[2] = (V)(-div3_rn(se));                  // (i-1, j+1)
    w[3] = (V)(-0.5 * div3_rn(sw + nw));       // (i, j-1)
    w[4] = (V)div3_rn(2.0 * (sw + se + nw + ne));
    w[5] = (V)(-0.5 * div3_rn(se + ne));       // (i, j+1)
    w[6] = (V)(-div3_rn(nw));                  // (i+1, j-1)
    w[7] = (V)(-0.5 * div3_rn(nw + ne));       // (i+1, j)
    w[8] = (V)(-div3_rn(ne));                  // (i+1, j+1)
}

// Node (i, j)'s cells read from coeff[(side-1)^2] (cell indices below 2^32:
// 32-bit arithmetic) and its weights formed.
template <typename V>
__device__ inline void q1_coeff_weights(int64_t side, int64_t i, int64_t j, const double *coeff, V (&w)[9]) {
    const uint32_t nc = (uint32_t)side - 1u, b = ((uint32_t)i - 1u) * nc + ((uint32_t)j - 1u);
    const double *cs = coeff + b, *cs2 = cs + nc;   // two adjacent pairs: two vector loads
    q1_cell_weights<V>(cs[0], cs[1], cs2[0], cs2[1], w);
}

// A float as a bf16 matrix stores it (round to nearest even) and widens it
// again on load.
__device__ inline float bf16_round(float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float((u >> 16) << 16);
}

// The weights of node (i, j) as a kernel reading the stored matrix of value
// type VT sees them (arithmetic type A: float for float / bf16 values,
// double for double): interior nodes from the cell coefficients, boundary
// nodes the identity row (w[0] = 1, the rest unused).
template <typename VT, typename A>
__device__ inline void q1_stored_weights(int64_t side, int64_t i, int64_t j, bool boundary, const double *coeff,
                                         A (&w)[9]) {
    if (boundary) {
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = A(0);
        w[0] = A(1);
        return;
    }
    q1_coeff_weights<A>(side, i, j, coeff, w);
    if constexpr (sizeof(VT) == 2) {
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = bf16_round(w[k]);
    }
}

// q1_stored_weights from the node's four cell values, already loaded.
template <typename VT, typename A>
__device__ inline void q1_stored_weights_cells(bool boundary, double sw, double se, double nw, double ne, A (&w)[9]) {
    if (boundary) {
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = A(0);
        w[0] = A(1);
        return;
    }
    q1_cell_weights<A>(sw, se, nw, ne, w);
    if constexpr (sizeof(VT) == 2) {
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = bf16_round(w[k]);
    }
}

// The interior weights from the node's four cell values, already loaded, as
// a kernel reading the stored matrix of value type VT sees them (no boundary
// case: a caller that forms every node's interior sum selects the identity
// row's sum for boundary nodes itself).
template <typename VT, typename A>
__device__ inline void q1_interior_weights_cells(double sw, double se, double nw, double ne, A (&w)[9]) {
    q1_cell_weights<A>(sw, se, nw, ne, w);
    if constexpr (sizeof(VT) == 2) {
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = bf16_round(w[k]);
    }
}

// Row `row` of the Q1 operator on a side x side grid in the T-interleaved
// ELL layout: entry t at slot (t % T) + row*T + (t / T)*stride, ncpr =
// ceil(9 / T) chunks, unused slots (and padding rows n <= row < stride/T)
// Aj = 0, Ax = 0.  Boundary rows: identity.  Inner rows: the constant
// stencil (coeff null: DT(8/3), DT(-1/3), as FillMatrix<POISSON,
// DIRICHLET_0>, math/fill_matrix.hh:36-78) or the bilinear element
// stiffness summed over the four cells around the node (coeff[(side-1)^2]).
// Aj / Arl null: values only.  dinv (optional) = omega / diag as stored.
// Entries are written at compile-time indices (no runtime-indexed register
// arrays).
template <typename VT>
__device__ inline void q1_row(int64_t side, int T, int64_t row, int32_t *Aj, VT *Ax, uint32_t *Arl, int64_t stride,
                              const double *coeff, typename VecOf<VT>::type *dinv, double omega) {
    typedef typename VecOf<VT>::type V;
    const int64_t n = side * side;
    const int ncpr = (9 + T - 1) / T;
    int cnt = 0;
    int64_t cols[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    V w[9] = {V(0), V(0), V(0), V(0), V(0), V(0), V(0), V(0), V(0)};
    if (row < n) {
        const int64_t i = row / side, j = row % side;
        const bool boundary = (i == 0 || i == side - 1 || j == 0 || j == side - 1);
        if (boundary) {
            cnt = 1;
            cols[0] = row;
            w[0] = V(1.0);
        } else {
            cnt = 9;
            if (!coeff) {
                const V c = V(-1.0 / 3.0);
                for (int k = 0; k < 9; ++k) w[k] = c;
                w[4] = V(8.0 / 3.0);
            } else {
                q1_coeff_weights<V>(side, i, j, coeff, w);
            }
            const int64_t c0[9] = {row - side - 1, row - side, row - side + 1, row - 1, row,
                                   row + 1,        row + side - 1, row + side, row + side + 1};
            for (int k = 0; k < 9; ++k) cols[k] = c0[k];
        }
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int64_t slot = (t % T) + row * T + (int64_t)(t / T) * stride;
        const bool have = t < cnt;
        if (Aj) Aj[slot] = have ? (int32_t)cols[t] : 0;
        q1_put<VT, V>(Ax, slot, have ? w[t] : V(0));
    }
    for (int t = 9; t < T * ncpr; ++t) {   // padding lanes of the last chunk
        const int64_t slot = (t % T) + row * T + (int64_t)(t / T) * stride;
        if (Aj) Aj[slot] = 0;
        q1_put<VT, V>(Ax, slot, V(0));
    }
    if (row >= n) return;
    if (Arl) Arl[row] = (uint32_t)((cnt + T - 1) / T);
    if (dinv) {
        V d = (cnt == 1) ? w[0] : w[4];
        if constexpr (sizeof(VT) == 2) d = bf16_round(d);   // diagonal as stored
        const V inv = V(1) / d;
        dinv[row] = inv * (V)omega;
    }
}

// The distinct weights of the variable-coefficient operator, one per cell,
// edge and node, interleaved per node (entry k of node r at st[4 r + k], so
// one 4-wide load fetches a node's four; unused entries untouched):
//   k = 0: -c / 3 of cell (i, j) at i * side + j (i, j < side - 1), the
//          corner weights w0, w2, w6, w8 of the cell's four nodes;
//   k = 1: -(c(i, j-1) + c(i, j)) / 6 of cell row i at i * side + j
//          (1 <= j <= side - 2): w1 of node (i + 1, j), w7 of node (i, j);
//   k = 2: -(c(i-1, j) + c(i, j)) / 6 of cell column j at i * side + j
//          (1 <= i <= side - 2): w3 of node (i, j + 1), w5 of node (i, j);
//   k = 3: the diagonal 2 (sw + se + nw + ne) / 3 of interior node (i, j).
// Each is the value q1_coeff_weights forms for every node that uses it (same
// operands, same order), rounded to the stored type as q1_row stores it, so
// the nine weights gathered from these arrays are the stored row's bits.
// Written for node index `node` < side^2.
template <typename VT>
__device__ inline void q1_stencil_node(int64_t side, int64_t node, const double *coeff, VT *st) {
    typedef typename VecOf<VT>::type V;
    const int64_t nc = side - 1, i = node / side, j = node % side;
    auto put = [&](int k, double v) {
        q1_put<VT, V>(st, 4 * node + k, (V)v);
    };
    if (i < nc && j < nc) put(0, -div3_rn(coeff[i * nc + j]));
    if (i < nc && j >= 1 && j <= side - 2) put(1, -0.5 * div3_rn(coeff[i * nc + (j - 1)] + coeff[i * nc + j]));
    if (j < nc && i >= 1 && i <= side - 2) put(2, -0.5 * div3_rn(coeff[(i - 1) * nc + j] + coeff[i * nc + j]));
    if (i >= 1 && i <= side - 2 && j >= 1 && j <= side - 2) {
        const double sw = coeff[(i - 1) * nc + (j - 1)], se = coeff[(i - 1) * nc + j];
        const double nw = coeff[i * nc + (j - 1)], ne = coeff[i * nc + j];
        put(3, div3_rn(2.0 * (sw + se + nw + ne)));
    }
}

// dinv = omega / diag of row `row` exactly as q1_row stores it, for a level
// whose values are not stored (interior rows: the coefficient operator).
template <typename VT>
__device__ inline void q1_dinv_row(int64_t side, int64_t row, const double *coeff, typename VecOf<VT>::type *dinv,
                                   double omega) {
    typedef typename VecOf<VT>::type V;
    const int64_t i = row / side, j = row % side;
    const bool boundary = (i == 0 || i == side - 1 || j == 0 || j == side - 1);
    V d = V(1.0);
    if (!boundary) {
        V w[9];
        q1_coeff_weights<V>(side, i, j, coeff, w);
        d = w[4];
    }
    if constexpr (sizeof(VT) == 2) d = bf16_round(d);   // diagonal as stored
    const V inv = V(1) / d;
    dinv[row] = inv * (V)omega;
}

// Coarse cell id = (I, J) of an nc x nc cell grid from fine cells (2I..2I+1,
// 2J..2J+1) of an nf = 2 nc grid: averaging 0 arithmetic, 1 geometric,
// 2 harmonic mean.
__device__ inline void coarsen_cell(const double *fine, int64_t nf, double *coarse, int64_t nc, int averaging,
                                    int64_t id) {
    const int64_t I = id / nc, J = id % nc;
    const double a = fine[(2 * I) * nf + 2 * J], b = fine[(2 * I) * nf + 2 * J + 1];
    const double c = fine[(2 * I + 1) * nf + 2 * J], d = fine[(2 * I + 1) * nf + 2 * J + 1];
    double m;
    if (averaging == 1)
        m = exp(0.25 * (((log(a) + log(b)) + log(c)) + log(d)));
    else if (averaging == 2)
        m = 4.0 / (((1.0 / a + 1.0 / b) + 1.0 / c) + 1.0 / d);
    else
        m = 0.25 * (((a + b) + c) + d);
    coarse[id] = m;
}

// Output number i+1 of the SplitMix64 sequence started at `seed` (the same
// draw the Python helper splitmix64(seed, n)[i] gives); U = (u >> 11) 2^-53.
__device__ inline uint64_t splitmix_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ inline double u01_at(uint64_t seed, uint64_t i) {
    return (double)(splitmix_at(seed, i) >> 11) * (1.0 / 9007199254740992.0);
}
// lo + (hi - lo) U_i
__device__ inline double uniform_at(double lo, double hi, uint64_t seed, uint64_t i) {
    return lo + (hi - lo) * u01_at(seed, i);
}
// exp(sigma z_i), z_i standard normal by Box-Muller from draws 2i (u1 = 1 -
// U, never 0) and 2i+1
__device__ inline double lognormal_at(double sigma, uint64_t seed, uint64_t i) {
    const double u1 = 1.0 - u01_at(seed, 2 * i);
    const double u2 = u01_at(seed, 2 * i + 1);
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    return exp(sigma * z);
}

// Start vector of the MIXED driver: first grid row, last grid row and first
// grid column take b (math/multigrid.hh:2423-2440).
__device__ inline bool guess_node(int64_t i, int64_t n, int64_t side) {
    return i < side || i >= n - side || (i % side) == 0;
}

}  // namespace hipx
}  // namespace honei
