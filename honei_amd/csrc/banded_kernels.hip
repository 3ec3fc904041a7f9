// banded_kernels.hip -- product and defect of the banded Q1 matrix
// (BandedMatrixQx<Q1Type, DT>: nine bands LL, LD, LU, DL, DD, DU, UL, UD, UU
// of length n, band k holding the coefficient of row i towards column
// i + offset_k, offsets -m-1, -m, -m+1, -1, 0, 1, m-1, m, m+1).
//
// No index stream: per row 9 values plus the gathered x, so the FP32 product
// moves 9*4 + 4 + 4 bytes per row against ELL's 9*(4+4) + 8.  Each row is
// summed in the CPU order of Product<tags::CPU>::value(BandedMatrixQx, DV)
// (la/product.hh:371-400: DD first, then LL, LD, LU, DL, DU, UL, UD, UU,
// each guarded by the column range), with separate multiply and add, so the
// result is bit-identical to the CPU path.  The defect is rhs - product, as
// Defect<tags::CPU> forms it (math/defect.hh:194-199).  Band and vector
// streams are read once with non-temporal loads; x comes through L2, where
// the three grid rows a block touches are shared by neighbouring blocks of
// the same XCD.

#include "backend.hh"

#pragma clang fp contract(off)

namespace honei {
namespace hipx {
namespace {

struct Bands9 {
    const void *b[9];   // LL, LD, LU, DL, DD, DU, UL, UD, UU
};

template <typename V, bool DEFECT>
__global__ __launch_bounds__(256) void k_q1_banded(Bands9 bands, const V *__restrict__ x, const V *__restrict__ rhs,
                                                   V *__restrict__ y, int64_t n, int64_t m) {
    const int64_t nb = gridDim.x;
    const int64_t i = xcd_block(blockIdx.x, nb) * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    // all nine band values and gathered x first (column indices clamped into
    // [0, n) so every load is in bounds), then the guarded sums in the CPU
    // order; a select keeps a term out exactly where the CPU loop skips it
    const int64_t off[9] = {-m - 1, -m, -m + 1, -1, 0, 1, m - 1, m, m + 1};
    V a[9], xv[9];
    bool ok[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int64_t j = i + off[k];
        ok[k] = j >= 0 && j < n;
        a[k] = __builtin_nontemporal_load((const V *)bands.b[k] + i);
        xv[k] = x[ok[k] ? j : i];
    }
    V r = a[4] * xv[4];                                             // DD
#pragma unroll
    for (int k = 0; k < 9; ++k)                                     // LL LD LU DL DU UL UD UU
        if (k != 4) r = ok[k] ? r + a[k] * xv[k] : r;
    if constexpr (DEFECT)
        y[i] = __builtin_nontemporal_load(rhs + i) - r;
    else
        y[i] = r;
}

}  // namespace

hipError_t q1_banded(ValueType vt, const void *const *bands, const void *x, const void *rhs, void *y, int64_t n,
                     int64_t m, int block, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (block <= 0) block = 256;
    if (block % kWave || block > 256 || m < 1 || m * m != n) return hipErrorInvalidValue;
    Bands9 b;
    for (int k = 0; k < 9; ++k) {
        if (!bands[k]) return hipErrorInvalidValue;
        b.b[k] = bands[k];
    }
    const int64_t grid = (n + block - 1) / block;
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    if (vt == VT_F32) {
        if (rhs)
            hipLaunchKernelGGL((k_q1_banded<float, true>), dim3((unsigned)grid), dim3(block), 0, s, b,
                               (const float *)x, (const float *)rhs, (float *)y, n, m);
        else
            hipLaunchKernelGGL((k_q1_banded<float, false>), dim3((unsigned)grid), dim3(block), 0, s, b,
                               (const float *)x, (const float *)nullptr, (float *)y, n, m);
    } else if (vt == VT_F64) {
        if (rhs)
            hipLaunchKernelGGL((k_q1_banded<double, true>), dim3((unsigned)grid), dim3(block), 0, s, b,
                               (const double *)x, (const double *)rhs, (double *)y, n, m);
        else
            hipLaunchKernelGGL((k_q1_banded<double, false>), dim3((unsigned)grid), dim3(block), 0, s, b,
                               (const double *)x, (const double *)nullptr, (double *)y, n, m);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace hipx
}  // namespace honei
