// solver_device.hh -- device pieces shared by the single-system and the
// batched solver kernels.
#pragma once

// Multiply and add stay separate roundings in everything that follows (the
// row sums and transfers reproduce the CPU path bit for bit).
#pragma clang fp contract(off)

#include "backend.hh"

namespace honei {
namespace hipx {

template <typename V> __device__ inline V eps_of();
template <> __device__ inline float eps_of<float>() { return 1.1920928955078125e-07f; }
template <> __device__ inline double eps_of<double>() { return 2.220446049250313e-16; }

// ---------------------------------------------------------------------------
// Single-work-group CG with the konv_rad stopping rule.  T = 1 layout or any
// T (row sums are formed serially per row, k-major then lane, like the CPU).
// ---------------------------------------------------------------------------
template <typename VT>
__device__ inline typename VecOf<VT>::type serial_row(const EllView &A, int64_t row,
                                                      const typename VecOf<VT>::type *x) {
    typedef typename VecOf<VT>::type V;
    const VT *Ax = (const VT *)A.Ax;
    V s = V(0);
    const uint32_t len = A.Arl[row];
    for (uint32_t k = 0; k < len; ++k)
        for (int t = 0; t < A.threads; ++t) {
            const int64_t slot = row * A.threads + t + (int64_t)k * A.stride;
            V av;
            if constexpr (sizeof(VT) == 2)
                av = bf16_to_float(((const uint16_t *)Ax)[slot]);
            else
                av = Ax[slot];
            s += av * x[A.Aj[slot]];
        }
    return s;
}

// Body of the coarse solve, run by every thread of one work-group; returns
// the iteration count.
template <typename VT>
__device__ inline int coarse_cg_block(const EllView &A, const typename VecOf<VT>::type *rhs,
                                      typename VecOf<VT>::type *x, typename VecOf<VT>::type *g,
                                      typename VecOf<VT>::type *u, typename VecOf<VT>::type *en, double konv,
                                      int max_iters) {
    typedef typename VecOf<VT>::type V;
    const int64_t n = A.rows;
    const V e = eps_of<V>();
    auto vnorm = [&](const V *a) -> V {
        double acc = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += (double)a[i] * (double)a[i];
        return (V)block_sum(acc);
    };
    auto vdot = [&](const V *a, const V *b) -> V {
        double acc = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += (double)a[i] * (double)b[i];
        return (V)block_sum(acc);
    };
    // x = 0 ; g = A x - rhs ; u = -g
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) x[i] = V(0);
    __syncthreads();
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const V gi = serial_row<VT>(A, i, x) - rhs[i];
        g[i] = gi;
        u[i] = gi * V(-1.);
    }
    __syncthreads();
    V norm_x = vnorm(x);
    V norm_x_last = V(norm_x - V(1));
    int it = 0;
    while ((double)(norm_x - norm_x_last) > konv && it < max_iters) {
        V upper = vnorm(g);
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) en[i] = serial_row<VT>(A, i, u);
        __syncthreads();
        V lower = vdot(en, u);
        const V f1 = (lower < 0 ? -lower : lower) >= e ? V(upper / lower) : V(upper / e);
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) en[i] = u[i] * f1 + x[i];
        __syncthreads();
        lower = upper;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) g[i] = serial_row<VT>(A, i, en) - rhs[i];
        __syncthreads();
        upper = vnorm(g);
        const V f2 = (lower < 0 ? -lower : lower) >= e ? V(upper / lower) : V(upper / e);
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            u[i] = u[i] * f2 - g[i];
            x[i] = en[i];
        }
        __syncthreads();
        ++it;
        norm_x_last = norm_x;
        norm_x = vnorm(x);
    }
    return it;
}

}  // namespace hipx
}  // namespace honei
