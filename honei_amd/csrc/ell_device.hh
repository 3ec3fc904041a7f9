// ell_device.hh -- device-side ELL row sum shared by the row kernels.
#pragma once

// Multiply and add stay separate roundings in everything that follows (the
// row sums and transfers reproduce the CPU path bit for bit).
#pragma clang fp contract(off)

#include "hip_common.hh"

namespace honei {
namespace hipx {

// Partial row sum of lane `lane` (0 <= lane < T) of row `row`: the lane walks
// slots row*T + lane + k*stride for k < len and adds Ax * x[Aj].  With
// SCALED the gathered operand is dinv[j] * x[j] instead (first Jacobi sweep
// from a zero start, where the iterate is dinv .* rhs).  Multiply and add are
// separate roundings (callers compile with fp contract off).  Ax is streamed
// with non-temporal loads; Aj too unless SHARED_AJ (one index array serving
// many value arrays, as in a batch, where it should stay in L2 / MALL).
// IW: 32-bit words per stored index -- 1 for this backend's int32 arrays, 2
// for the reference's 64-bit unsigned long arrays (la/sparse_matrix_ell.hh:
// 110-116), of which the low word is read (little-endian; column indices
// below 2^31).
template <typename VT, int T, bool SCALED = false, bool SHARED_AJ = false, int IW = 1>
__device__ inline typename VecOf<VT>::type ell_lane_sum(const int32_t *__restrict__ Aj, const VT *__restrict__ Ax,
                                                        int64_t row, int lane, uint32_t len, int64_t stride,
                                                        const typename VecOf<VT>::type *__restrict__ x,
                                                        const typename VecOf<VT>::type *__restrict__ dinv = nullptr) {
    typedef typename VecOf<VT>::type V;
    V s = V(0);
    int64_t slot = row * T + lane;
#pragma unroll 4
    for (uint32_t k = 0; k < len; ++k) {
        int32_t j;
        if constexpr (SHARED_AJ)
            j = Aj[slot * IW];
        else
            j = __builtin_nontemporal_load(Aj + slot * IW);
        const V av = load_value_nt(Ax + slot);
        V xv;
        if constexpr (SCALED)
            xv = dinv[j] * x[j];
        else
            xv = x[j];
        s += av * xv;
        slot += stride;
    }
    return s;
}

// ell_lane_sum for G value arrays that share one index array (G systems of
// one sparsity pattern) whose stored width ncpr is the compile-time NC: the
// lane's NC column indices are loaded once for all G, then all values and
// gathered x, so a row costs two dependent memory round trips instead of one
// pair per unrolled step.  Slots k >= len are read too (padding holds Aj =
// 0, Ax = 0, inside the ncpr*stride arrays) but never added: per array each
// sum is the same sequence of roundings as ell_lane_sum.  Every array is
// read (a caller with fewer live systems points the spare entries at a live
// one and ignores their sums).
template <typename VT, int T, int G, int NC, bool SCALED = false>
__device__ inline void ell_lane_sum_group_w(const int32_t *__restrict__ Aj, const VT *const (&Ax)[G], int64_t row,
                                            int lane, uint32_t len, int64_t stride,
                                            const typename VecOf<VT>::type *const (&x)[G],
                                            const typename VecOf<VT>::type *const (&dinv)[G],
                                            typename VecOf<VT>::type (&s)[G]) {
    typedef typename VecOf<VT>::type V;
    const int64_t slot0 = row * T + lane;
    int32_t j[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) j[k] = Aj[slot0 + k * stride];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        V av[NC], xv[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            av[k] = load_value_nt(Ax[g] + slot0 + k * stride);
            if constexpr (SCALED)
                xv[k] = dinv[g][j[k]] * x[g][j[k]];
            else
                xv[k] = x[g][j[k]];
        }
        V t = V(0);
#pragma unroll
        for (int k = 0; k < NC; ++k) t = ((uint32_t)k < len) ? t + av[k] * xv[k] : t;
        s[g] = t;
    }
}

// ---- buffer loads -----------------------------------------------------------
// A buffer resource over [p, p + bytes) (raw, 32-bit offsets; reads past the
// end return 0).  Loads through it take a per-lane 32-bit byte offset plus a
// uniform scalar one, so the NC slots of a row (slot0 + k*stride) share one
// offset register and differ only in the scalar k*stride*size.
__device__ inline __amdgpu_buffer_rsrc_t buffer_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
constexpr int kBufNT = 2;   // cache policy bit: non-temporal

template <typename T> struct BufLoad;
template <> struct BufLoad<float> {
    template <int AUX> static __device__ inline float load(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, AUX));
    }
};
template <> struct BufLoad<double> {
    template <int AUX> static __device__ inline double load(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, AUX));
    }
};
template <> struct BufLoad<int32_t> {
    template <int AUX> static __device__ inline int32_t load(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
        return (int32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, AUX);
    }
};
template <> struct BufLoad<bf16> {   // the bf16 value widened to float
    template <int AUX> static __device__ inline float load(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
        const uint32_t h = __builtin_amdgcn_raw_buffer_load_b16(r, (int)vo, (int)so, AUX);
        return __uint_as_float(h << 16);
    }
};

// ell_lane_sum for a matrix of stored width NC whose index and value
// streams are read through buffer resources (rAj over Aj, rAx over Ax): the
// NC indices first, then the NC values and the gathered x (plain loads: the
// length of x is not known here), then the guarded sums -- each sum is the
// same sequence of roundings as ell_lane_sum (slots k >= len are read,
// padding holds Aj = 0 and Ax = 0, but never added).  NC*stride slots must
// fit 31-bit byte offsets (the caller checks).  IW as for ell_lane_sum.
template <typename VT, int T, int NC, bool SCALED = false, int IW = 1>
__device__ inline typename VecOf<VT>::type ell_lane_sum_buf(__amdgpu_buffer_rsrc_t rAj, __amdgpu_buffer_rsrc_t rAx,
                                                            int64_t row, int lane, uint32_t len, int64_t stride,
                                                            const typename VecOf<VT>::type *__restrict__ x,
                                                            const typename VecOf<VT>::type *__restrict__ dinv) {
    typedef typename VecOf<VT>::type V;
    const uint32_t slot0 = (uint32_t)(row * T + lane);
    int32_t j[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k)
        j[k] = BufLoad<int32_t>::template load<kBufNT>(rAj, slot0 * (4u * IW), (uint32_t)(k * stride) * (4u * IW));
    V av[NC], xv[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        av[k] = BufLoad<VT>::template load<kBufNT>(rAx, slot0 * (uint32_t)sizeof(VT),
                                                   (uint32_t)(k * stride) * (uint32_t)sizeof(VT));
        if constexpr (SCALED)
            xv[k] = dinv[j[k]] * x[j[k]];
        else
            xv[k] = x[j[k]];
    }
    V s = V(0);
#pragma unroll
    for (int k = 0; k < NC; ++k) s = ((uint32_t)k < len) ? s + av[k] * xv[k] : s;
    return s;
}

// Stored width of the Q1 operator (9 entries per row) at T lanes per row.
constexpr int q1_width(int T) { return (9 + T - 1) / T; }

template <int T, typename V> __device__ inline V lane_group_sum(V s) {
    if constexpr (T > 1) {
#pragma unroll
        for (int off = T / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
    }
    return s;
}

}  // namespace hipx
}  // namespace honei
