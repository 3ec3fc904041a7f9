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

template <typename T> __device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Block-wide sum (any blockDim multiple of 64, at most 1024); result valid in
// every thread.  Fixed order, hence deterministic.
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
