// hip_common.hh -- shared device/host helpers of the MI355X (gfx950) backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace honei {
namespace hipx {

constexpr int kWave = 64;       // CDNA wavefront width
constexpr int kXcds = 8;        // MI355X: 8 XCDs, each with its own L2

// bf16 storage (matrix values only; arithmetic is FP32).
struct bf16 {
    uint16_t bits;
};

__host__ __device__ inline float to_float(float v) { return v; }
__host__ __device__ inline double to_double(double v) { return v; }
__device__ inline float bf16_to_float(uint16_t b) {
    return __uint_as_float(((uint32_t)b) << 16);
}

// Vector element type for a matrix value type (bf16 matrices act on f32).
template <typename VT> struct VecOf { typedef VT type; };
template <> struct VecOf<bf16> { typedef float type; };

// Accumulation type of a value type.
template <typename VT> struct AccOf { typedef VT type; };
template <> struct AccOf<bf16> { typedef float type; };

template <typename VT> __device__ inline typename AccOf<VT>::type load_value_nt(const VT *p) {
    return __builtin_nontemporal_load(p);
}
template <> __device__ inline float load_value_nt<bf16>(const bf16 *p) {
    return bf16_to_float(__builtin_nontemporal_load(&p->bits));
}

// The hardware hands work-group b to XCD b % 8.  Give every XCD one
// contiguous range of logical blocks instead, so rows that share x entries
// (neighbouring grid rows) meet in the same L2.  Bijective for any count.
__device__ inline int64_t xcd_block(int64_t b, int64_t nblocks) {
    const int64_t q = nblocks / kXcds, r = nblocks % kXcds;
    const int64_t x = b % kXcds, i = b / kXcds;
    return x * q + (x < r ? x : r) + i;
}

// Inverse of xcd_block: the b with xcd_block(b, nblocks) == rb.
__device__ inline int64_t xcd_unblock(int64_t rb, int64_t nblocks) {
    const int64_t q = nblocks / kXcds, r = nblocks % kXcds;
    int64_t x, i;
    if (rb < r * (q + 1)) {
        x = rb / (q + 1);
        i = rb - x * (q + 1);
    } else {
        const int64_t t = rb - r * (q + 1);
        x = r + t / q;
        i = t - (x - r) * q;
    }
    return i * kXcds + x;
}

// Lane l < off of `v` receives lane l + off's value (other lanes: unspecified):
// off 32 / 16 through the gfx950 lane-swap permutes, off 8 .. 1 through DPP
// row shifts, every 32-bit half on its own -- register moves, no LDS trip.
template <int OFF> __device__ inline uint32_t lane_down32(uint32_t v) {
    if constexpr (OFF == 32) {
        return __builtin_amdgcn_permlane32_swap(v, v, false, false)[1];   // src lanes 0..31 <- lanes 32..63
    } else if constexpr (OFF == 16) {
        return __builtin_amdgcn_permlane16_swap(v, v, false, false)[1];   // src row 0 <- row 1, row 2 <- row 3
    } else {
        static_assert(OFF >= 1 && OFF <= 8, "row shift");
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x100 + OFF, 0xF, 0xF, true);   // row_shl:OFF
    }
}
template <int OFF, typename T> __device__ inline T lane_down(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit lanes");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, lane_down32<OFF>(__builtin_bit_cast(uint32_t, v)));
    } else {
        const uint64_t b = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = lane_down32<OFF>((uint32_t)b), hi = lane_down32<OFF>((uint32_t)(b >> 32));
        return __builtin_bit_cast(T, lo | (hi << 32));
    }
}

// Sum of a wavefront's values, in lane 0 (other lanes: partial sums).  Lane
// 0 adds the pairs of the xor butterfly (lanes l and l + 32, then + 16, ...
// + 1), so its bits are those of that tree.  Every lane of the wavefront
// must be active.
template <typename T> __device__ inline T wave_sum(T v) {
    v += lane_down<32>(v);
    v += lane_down<16>(v);
    v += lane_down<8>(v);
    v += lane_down<4>(v);
    v += lane_down<2>(v);
    v += lane_down<1>(v);
    return v;
}

// Block-wide sum (any blockDim multiple of 64, at most 1024); result valid in
// every thread (wave_sum's lane 0 values, broadcast through LDS).  Fixed order, hence deterministic.
template <typename T> __device__ inline T block_sum(T v) {
    __shared__ T red[1024 / kWave];
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    const int nw = blockDim.x / kWave;
    if (l == 0) red[w] = v;
    __syncthreads();
    T t = (threadIdx.x < (unsigned)nw) ? red[threadIdx.x] : T(0);
    if (w == 0) t = wave_sum(t);
    if (threadIdx.x == 0) red[0] = t;
    __syncthreads();
    t = red[0];
    __syncthreads();
    return t;
}

// G block-wide sums at once, each in block_sum's order (so bit-identical to
// it); the results are valid in thread 0 only.  Ends with a barrier, so the
// next call may follow at once.
template <int G, typename T> __device__ inline void block_sum_first(T (&v)[G]) {
    __shared__ T red[G][1024 / kWave];
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    const int nw = blockDim.x / kWave;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const T s = wave_sum(v[k]);
        if (l == 0) red[k][w] = s;
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const T t = (threadIdx.x < (unsigned)nw) ? red[k][threadIdx.x] : T(0);
            v[k] = wave_sum(t);
        }
    }
    __syncthreads();
}

// A flag read once per work-group (thread 0, broadcast through LDS): every
// wavefront of the work-group takes the same branch on it even while another
// work-group of the same launch is writing the flag, so no wavefront leaves
// early while the others go on to a barrier or a block-wide sum.
__device__ inline int block_uniform_flag(const int *flag) {
    __shared__ int f;
    if (threadIdx.x == 0) f = *(const volatile int *)flag;
    __syncthreads();
    const int v = f;
    __syncthreads();
    return v;
}

// Cross-block reduction slots.  A reducing kernel stores one partial per
// block with a plain store; a one-block finisher launched behind it sums the
// partials in a fixed order.  (A last-block-arrives scheme would need an
// agent-scope release per block, which on MI355X writes back the XCD's L2 --
// far more expensive than the extra launch.)
struct ReduceSlots {
    double *partials;      // [capacity][lanes]
    unsigned int *counter; // unused by the partial/finisher scheme; kept zero
    int capacity;          // max blocks
};

// Deterministic sum of partials[b*lanes + lane] over b < nparts by one block.
__device__ inline double sum_partials(const double *partials, int64_t nparts, int lanes, int lane) {
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < nparts; b += blockDim.x) v += partials[b * lanes + lane];
    return block_sum(v);
}

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace hipx
}  // namespace honei
