// transfer_kernels.hip -- geometric multigrid transfers on lexicographic
// (2^L+1)^2 Q1 grids, one thread per output node (node arithmetic in
// transfer_device.hh).

#include "backend.hh"
#include "transfer_device.hh"

#pragma clang fp contract(off)

namespace honei {
namespace hipx {
namespace {

template <typename V>
__global__ __launch_bounds__(256) void k_restriction(V *__restrict__ coarse, const V *__restrict__ fine, int64_t m, Mask8 mk) {
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id < (m + 1) * (m + 1)) restrict_node(coarse, fine, m, mk, id);
}

// fine[R,C] from coarse, in the reference's accumulation order; optionally
// x[R,C] += alpha * fine[R,C] (the MG correction ScaledSum, fused).
template <typename V, bool AXPY>
__global__ __launch_bounds__(256) void k_prolongation(V *__restrict__ fine, const V *__restrict__ coarse, int64_t m, Mask8 mk,
                               V *__restrict__ x, V alpha) {
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id < (2 * m + 1) * (2 * m + 1)) prolong_node<V, AXPY>(fine, coarse, m, mk, x, alpha, id);
}

bool side_of(int64_t n, int64_t &side) {
    int64_t s = 1;
    while (s * s < n) ++s;
    side = s;
    return s * s == n && s >= 3 && s <= kMaxTransferSide && ((s - 1) & (s - 2)) == 0;
}


}  // namespace

hipError_t restriction(void *coarse, int64_t n_coarse, const void *fine, int64_t n_fine, const uint64_t *mask,
                       ValueType vt, hipStream_t s) {
    int64_t sf, sc;
    if (!side_of(n_fine, sf) || !side_of(n_coarse, sc) || sc != (sf - 1) / 2 + 1) return hipErrorInvalidValue;
    const int64_t m = sc - 1;
    const int grid = ceil_div(n_coarse, 256);
    if (vt == VT_F32)
        hipLaunchKernelGGL(k_restriction<float>, dim3(grid), dim3(256), 0, s, (float *)coarse, (const float *)fine, m,
                           pack_mask(mask));
    else if (vt == VT_F64)
        hipLaunchKernelGGL(k_restriction<double>, dim3(grid), dim3(256), 0, s, (double *)coarse, (const double *)fine,
                           m, pack_mask(mask));
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t prolongation_axpy(void *fine, int64_t n_fine, const void *coarse, int64_t n_coarse, const uint64_t *mask,
                             void *x, double alpha, ValueType vt, hipStream_t s) {
    int64_t sf, sc;
    if (!side_of(n_fine, sf) || !side_of(n_coarse, sc) || sc != (sf - 1) / 2 + 1) return hipErrorInvalidValue;
    const int64_t m = sc - 1;
    const int grid = ceil_div(n_fine, 256);
    if (vt == VT_F32) {
        if (x)
            hipLaunchKernelGGL((k_prolongation<float, true>), dim3(grid), dim3(256), 0, s, (float *)fine,
                               (const float *)coarse, m, pack_mask(mask), (float *)x, (float)alpha);
        else
            hipLaunchKernelGGL((k_prolongation<float, false>), dim3(grid), dim3(256), 0, s, (float *)fine,
                               (const float *)coarse, m, pack_mask(mask), (float *)nullptr, 0.f);
    } else if (vt == VT_F64) {
        if (x)
            hipLaunchKernelGGL((k_prolongation<double, true>), dim3(grid), dim3(256), 0, s, (double *)fine,
                               (const double *)coarse, m, pack_mask(mask), (double *)x, alpha);
        else
            hipLaunchKernelGGL((k_prolongation<double, false>), dim3(grid), dim3(256), 0, s, (double *)fine,
                               (const double *)coarse, m, pack_mask(mask), (double *)nullptr, 0.0);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t prolongation(void *fine, int64_t n_fine, const void *coarse, int64_t n_coarse, const uint64_t *mask,
                        ValueType vt, hipStream_t s) {
    return prolongation_axpy(fine, n_fine, coarse, n_coarse, mask, nullptr, 0.0, vt, s);
}

}  // namespace hipx
}  // namespace honei
