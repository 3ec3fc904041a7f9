// blas1_kernels.hip -- streaming vector kernels and single-pass reductions.
//
// Elementwise kernels are grid-stride loops over 64-wide wavefronts.
// Reductions accumulate in double, combine per block in a fixed tree, store
// one partial per block, and a one-block finisher forms the total on the
// device (no host round trip, no per-call allocation -- unlike
// backends/cuda/dot_product.cu:95-109, which copies grid*block partials to the
// host).  The grid size depends only on n, so results are reproducible.

#include "backend.hh"

#pragma clang fp contract(off)

namespace honei {
namespace hipx {
namespace {

constexpr int kBlock = 256;

inline int stream_grid(int64_t n) {
    int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return (int)g;
}
inline int reduce_grid(int64_t n) {
    int64_t g = (n + kBlock * 8 - 1) / (kBlock * 8);
    if (g < 1) g = 1;
    if (g > 2048) g = 2048;
    return (int)g;
}

#define GRID_STRIDE(i, n) \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

template <typename V> __global__ void k_axpy(V *x, const V *y, V b, const double *b_dev, int64_t n, const int *skip) {
    if (skip && block_uniform_flag(skip)) return;
    const V bb = b_dev ? (V)*b_dev : b;
    GRID_STRIDE(i, n) x[i] += y[i] * bb;
}
template <typename V> __global__ void k_scaled_sum3(V *r, const V *x, const V *y, V b, int64_t n) {
    GRID_STRIDE(i, n) r[i] = x[i] + y[i] * b;
}
template <typename V> __global__ void k_scaled_sum_v(V *a, const V *b, const V *c, int64_t n) {
    GRID_STRIDE(i, n) a[i] += b[i] * c[i];
}
template <typename V> __global__ void k_sum(V *x, const V *y, int64_t n) { GRID_STRIDE(i, n) x[i] += y[i]; }
template <typename V> __global__ void k_diff2(V *x, const V *y, int64_t n) { GRID_STRIDE(i, n) x[i] -= y[i]; }
template <typename V> __global__ void k_diff3(V *r, const V *x, const V *y, int64_t n) {
    GRID_STRIDE(i, n) r[i] = x[i] - y[i];
}
template <typename V> __global__ void k_scale(V *x, V a, int64_t n) { GRID_STRIDE(i, n) x[i] *= a; }
template <typename V> __global__ void k_prod(V *r, const V *x, const V *y, int64_t n) {
    GRID_STRIDE(i, n) r[i] = x[i] * y[i];
}
template <typename V> __global__ void k_inverse(V *x, int64_t n) { GRID_STRIDE(i, n) x[i] = V(1) / x[i]; }
template <typename V> __global__ void k_fill(V *x, V v, int64_t n) { GRID_STRIDE(i, n) x[i] = v; }
template <typename V> __global__ void k_copy(V *d, const V *s, int64_t n) { GRID_STRIDE(i, n) d[i] = s[i]; }
__global__ void k_d2f(float *d, const double *s, int64_t n) { GRID_STRIDE(i, n) d[i] = (float)s[i]; }
__global__ void k_f2d(double *d, const float *s, int64_t n) { GRID_STRIDE(i, n) d[i] = (double)s[i]; }
// d[i] = low word of s[i]; entries >= bound counted into *bad (one vector
// atomic per thread that met any)
__global__ void k_narrow_u64(uint32_t *d, const uint64_t *s, int64_t n, uint64_t bound, unsigned int *bad) {
    unsigned int nb = 0;
    GRID_STRIDE(i, n) {
        const uint64_t v = s[i];
        nb += v >= bound ? 1u : 0u;
        d[i] = (uint32_t)v;
    }
    if (nb) atomicAdd(bad, nb);
}
// Bandwidth probes: 16-byte non-temporal loads (and stores), eight in flight
// per thread, grid-stride over the array.  The read probe keeps a sum per
// work-group (written to `partials`) so the loads are not removed.
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int kProbeUnroll = 8;
__global__ __launch_bounds__(256) void k_probe_read(const f4v *__restrict__ src, int64_t n4, float *partials) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    for (; i + (kProbeUnroll - 1) * stride < n4; i += kProbeUnroll * stride) {
        f4v v[kProbeUnroll];
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) v[k] = __builtin_nontemporal_load(src + i + k * stride);
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) acc += v[k];
    }
    for (; i < n4; i += stride) acc += __builtin_nontemporal_load(src + i);
    const float t = block_sum(acc.x + acc.y + acc.z + acc.w);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}
__global__ __launch_bounds__(256) void k_probe_copy(f4v *__restrict__ dst, const f4v *__restrict__ src, int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i + (kProbeUnroll - 1) * stride < n4; i += kProbeUnroll * stride) {
        f4v v[kProbeUnroll];
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) v[k] = __builtin_nontemporal_load(src + i + k * stride);
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) __builtin_nontemporal_store(v[k], dst + i + k * stride);
    }
    for (; i < n4; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}
// Chunked forms: work-group b (remapped so each XCD takes one contiguous
// range of chunks, as the SpMV's row blocks) reads (and copies) one
// contiguous chunk of `per` 16-byte elements, eight loads in flight per
// thread; the grid covers the array once.
__global__ __launch_bounds__(256) void k_probe_read_chunked(const f4v *__restrict__ src, int64_t n4, int64_t per,
                                                            float *partials) {
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);
    const int64_t i0 = b * per, i1 = min(n4, i0 + per);
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    int64_t i = i0 + threadIdx.x;
    for (; i + (kProbeUnroll - 1) * 256 < i1; i += kProbeUnroll * 256) {
        f4v v[kProbeUnroll];
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) v[k] = __builtin_nontemporal_load(src + i + k * 256);
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) acc += v[k];
    }
    for (; i < i1; i += 256) acc += __builtin_nontemporal_load(src + i);
    const float t = block_sum(acc.x + acc.y + acc.z + acc.w);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}
__global__ __launch_bounds__(256) void k_probe_copy_chunked(f4v *__restrict__ dst, const f4v *__restrict__ src,
                                                            int64_t n4, int64_t per) {
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);
    const int64_t i0 = b * per, i1 = min(n4, i0 + per);
    int64_t i = i0 + threadIdx.x;
    for (; i + (kProbeUnroll - 1) * 256 < i1; i += kProbeUnroll * 256) {
        f4v v[kProbeUnroll];
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) v[k] = __builtin_nontemporal_load(src + i + k * 256);
#pragma unroll
        for (int k = 0; k < kProbeUnroll; ++k) __builtin_nontemporal_store(v[k], dst + i + k * 256);
    }
    for (; i < i1; i += 256) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}
__global__ void k_zero_bytes(unsigned char *p, size_t n) {
    GRID_STRIDE(i, (int64_t)n) p[i] = 0;
}
__global__ void k_zero_words(uint32_t *p, size_t n) {
    GRID_STRIDE(i, (int64_t)n) p[i] = 0u;
}

// MODE 0: sum x*y ; 1: sum x*x ; 2: x += a*y then sum x*x.  Each thread
// walks its grid-stride elements four at a time into four accumulators
// (the four loads in flight together), summed in a fixed order at the end;
// the grid depends on n only, so the result is reproducible.
template <typename V, int MODE>
__global__ __launch_bounds__(kBlock) void k_reduce(V *x, const V *y, V a, const double *a_dev, int64_t n,
                                                   double *out, ReduceSlots slots, const int *skip) {
    if (skip && block_uniform_flag(skip)) return;
    V aa = a;
    if constexpr (MODE == 2) {
        if (a_dev) aa = (V)*a_dev;
    }
    auto term = [&](int64_t i) -> double {
        if constexpr (MODE == 0) {
            return (double)x[i] * (double)y[i];
        } else if constexpr (MODE == 1) {
            return (double)x[i] * (double)x[i];
        } else {
            const V v = x[i] + y[i] * aa;
            x[i] = v;
            return (double)v * (double)v;
        }
    };
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (; i + 3 * stride < n; i += 4 * stride) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] += term(i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < 3; ++u)   // at most three left; compile-time accumulator index
        if (i + u * stride < n) acc[u] += term(i + u * stride);
    const double b = block_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
    if (threadIdx.x == 0) slots.partials[blockIdx.x] = b;
}

__global__ void k_finish_sum(const double *partials, int64_t nparts, double *out, const int *skip) {
    if (skip && block_uniform_flag(skip)) return;
    const double t = sum_partials(partials, nparts, 1, 0);
    if (threadIdx.x == 0) *out = t;
}

template <typename F>
hipError_t by_type(ValueType vt, F &&f) {
    if (vt == VT_F32) return f(float(0));
    if (vt == VT_F64) return f(double(0));
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t dot(const void *x, const void *y, int64_t n, ValueType vt, double *out, const ReduceSlots &sl,
               hipStream_t s) {
    return by_type(vt, [&](auto z) {
        typedef decltype(z) V;
        const int g = reduce_grid(n);
        hipLaunchKernelGGL((k_reduce<V, 0>), dim3(g), dim3(kBlock), 0, s, (V *)x, (const V *)y,
                           V(0), (const double *)nullptr, n, out, sl, (const int *)nullptr);
        hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, s, (const double *)sl.partials, (int64_t)g, out,
                           (const int *)nullptr);
        return hipGetLastError();
    });
}
hipError_t norm2(const void *x, int64_t n, ValueType vt, double *out, const ReduceSlots &sl, hipStream_t s) {
    return by_type(vt, [&](auto z) {
        typedef decltype(z) V;
        const int g = reduce_grid(n);
        hipLaunchKernelGGL((k_reduce<V, 1>), dim3(g), dim3(kBlock), 0, s, (V *)x, (const V *)x,
                           V(0), (const double *)nullptr, n, out, sl, (const int *)nullptr);
        hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, s, (const double *)sl.partials, (int64_t)g, out,
                           (const int *)nullptr);
        return hipGetLastError();
    });
}
hipError_t axpy_norm2(void *x, const void *y, double alpha, const double *alpha_dev, int64_t n, ValueType vt,
                      double *out, const ReduceSlots &sl, hipStream_t s, const int *skip) {
    return by_type(vt, [&](auto z) {
        typedef decltype(z) V;
        const int g = reduce_grid(n);
        hipLaunchKernelGGL((k_reduce<V, 2>), dim3(g), dim3(kBlock), 0, s, (V *)x, (const V *)y,
                           (V)alpha, alpha_dev, n, out, sl, skip);
        hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, s, (const double *)sl.partials, (int64_t)g, out,
                           skip);
        return hipGetLastError();
    });
}

#define ELEMENTWISE(launch)                                              \
    return by_type(vt, [&](auto z) {                                    \
        typedef decltype(z) V;                                           \
        (void)sizeof(V);                                                 \
        launch;                                                          \
        return hipGetLastError();                                        \
    })

hipError_t axpy(void *x, const void *y, double b, const double *b_dev, int64_t n, ValueType vt, hipStream_t s,
                const int *skip) {
    ELEMENTWISE(hipLaunchKernelGGL((k_axpy<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)x, (const V *)y,
                                   (V)b, b_dev, n, skip));
}
hipError_t scaled_sum3(void *r, const void *x, const void *y, double b, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_scaled_sum3<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)r,
                                   (const V *)x, (const V *)y, (V)b, n));
}
hipError_t scaled_sum_v(void *a, const void *b, const void *c, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_scaled_sum_v<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)a,
                                   (const V *)b, (const V *)c, n));
}
hipError_t vsum(void *x, const void *y, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_sum<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)x, (const V *)y, n));
}
hipError_t vdiff2(void *x, const void *y, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_diff2<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)x, (const V *)y, n));
}
hipError_t vdiff3(void *r, const void *x, const void *y, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_diff3<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)r, (const V *)x,
                                   (const V *)y, n));
}
hipError_t vscale(void *x, double a, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_scale<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)x, (V)a, n));
}
hipError_t vprod(void *r, const void *x, const void *y, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_prod<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)r, (const V *)x,
                                   (const V *)y, n));
}
hipError_t vinverse(void *x, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_inverse<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)x, n));
}
hipError_t vfill(void *x, double v, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_fill<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)x, (V)v, n));
}
hipError_t vcopy(void *d, const void *src, int64_t n, ValueType vt, hipStream_t s) {
    ELEMENTWISE(hipLaunchKernelGGL((k_copy<V>), dim3(stream_grid(n)), dim3(kBlock), 0, s, (V *)d,
                                   (const V *)src, n));
}
hipError_t convert_d2f(float *d, const double *src, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_d2f, dim3(stream_grid(n)), dim3(kBlock), 0, s, d, src, n);
    return hipGetLastError();
}
hipError_t convert_f2d(double *d, const float *src, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_f2d, dim3(stream_grid(n)), dim3(kBlock), 0, s, d, src, n);
    return hipGetLastError();
}
hipError_t narrow_indices(uint32_t *d, const uint64_t *src, int64_t n, uint64_t bound, unsigned int *bad,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_narrow_u64, dim3(stream_grid(n)), dim3(kBlock), 0, s, d, src, n, bound, bad);
    return hipGetLastError();
}
hipError_t stream_probe(int kind, void *dst, const void *src, int64_t n4, int grid, hipStream_t s) {
    const int64_t per = (n4 + grid - 1) / grid;   // chunked forms: elements per work-group
    switch (kind) {
        case 0:
            hipLaunchKernelGGL(k_probe_read, dim3(grid), dim3(kBlock), 0, s, (const f4v *)src, n4, (float *)dst);
            break;
        case 1:
            hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(kBlock), 0, s, (f4v *)dst, (const f4v *)src, n4);
            break;
        case 2:
            hipLaunchKernelGGL(k_probe_read_chunked, dim3(grid), dim3(256), 0, s, (const f4v *)src, n4, per,
                               (float *)dst);
            break;
        case 3:
            hipLaunchKernelGGL(k_probe_copy_chunked, dim3(grid), dim3(256), 0, s, (f4v *)dst, (const f4v *)src, n4,
                               per);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t fill_bytes_zero(void *p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    if (((uintptr_t)p % 4 == 0) && bytes % 4 == 0) {
        hipLaunchKernelGGL(k_zero_words, dim3(stream_grid((int64_t)bytes / 4)), dim3(kBlock), 0, s, (uint32_t *)p,
                           bytes / 4);
    } else {
        hipLaunchKernelGGL(k_zero_bytes, dim3(stream_grid((int64_t)bytes)), dim3(kBlock), 0, s,
                           (unsigned char *)p, bytes);
    }
    return hipGetLastError();
}

}  // namespace hipx
}  // namespace honei
