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
template <typename VT, int T, bool SCALED = false, bool SHARED_AJ = false>
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
            j = Aj[slot];
        else
            j = __builtin_nontemporal_load(Aj + slot);
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

template <int T, typename V> __device__ inline V lane_group_sum(V s) {
    if constexpr (T > 1) {
#pragma unroll
        for (int off = T / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
    }
    return s;
}

}  // namespace hipx
}  // namespace honei
