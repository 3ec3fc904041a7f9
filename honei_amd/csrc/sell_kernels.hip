// sell_kernels.hip -- sliced copy of an ELL matrix for rows of uneven length.
//
// The ELL layout pads every row to the widest row's length, and a product
// streams all num_cols_per_row * stride slots.  For matrices whose row lengths
// vary widely (the random 4M-row matrix of the SpMV benchmark: lengths
// uniform in 1..27, so half the slots are padding) this copy groups rows into
// slices of 64 (one wavefront, a lane per row) after sorting the rows of each
// window of 2048 by length, longest first, and pads each slice only to its
// own longest row.  Each row keeps its entries in their stored order, and the
// product adds them in that order with separate roundings, so y equals the
// ELL product of a one-lane-per-row (T = 1) matrix, and Product<tags::CPU>
// (la/product.hh:148-206), bit for bit; only the order in which rows are
// computed changes.
//
// Layout (device): slice s covers sorted positions 64 s .. 64 s + 63 and has
// width w_s = the longest row among them; entry k of the row at position
// 64 s + l sits at slot base_s + 64 k + l (base_s = 64 * sum of the earlier
// widths).  perm maps positions to rows, len holds each position's length.

#include "backend.hh"
#include "hip_common.hh"

#pragma clang fp contract(off)

namespace honei {
namespace hipx {
namespace {

constexpr int kSellC = 64;          // rows per slice
constexpr int kSellWindow = 2048;   // rows sorted together
constexpr int kSellKeys = 64;       // sort keys: lengths 0..62 and 63+ (longer rows share the last key)
constexpr int kSellChunks = kSellWindow / kSellC;

// perm / len of window blockIdx.x: rows by key descending, ties by row index
// (a counting sort; ranks from wavefront ballots, so the order is fixed).
__global__ __launch_bounds__(1024) void k_sell_sort(const uint32_t *Arl, int64_t rows, int32_t *perm, uint32_t *len) {
    __shared__ int cnt[kSellChunks][kSellKeys];
    __shared__ int tot[kSellKeys], start[kSellKeys];
    const int64_t w0 = (int64_t)blockIdx.x * kSellWindow;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int waves = blockDim.x / kWave;
    constexpr int kPer = 2;   // chunks per wavefront (1024 threads, 32 chunks)
    int key[kPer], rank[kPer];
    uint32_t ln[kPer];
    bool valid[kPer];
#pragma unroll
    for (int h = 0; h < kPer; ++h) {
        const int chunk = wave + waves * h;
        const int64_t row = w0 + (int64_t)chunk * kSellC + lane;
        valid[h] = chunk < kSellChunks && row < rows;
        ln[h] = valid[h] ? Arl[row] : 0u;
        key[h] = valid[h] ? (int)(ln[h] < (uint32_t)(kSellKeys - 1) ? ln[h] : (uint32_t)(kSellKeys - 1)) : -1;
        rank[h] = 0;
        for (int k = 0; k < kSellKeys; ++k) {
            const uint64_t b = __ballot(key[h] == k);
            if (chunk < kSellChunks && lane == 0) cnt[chunk][k] = __popcll(b);
            if (key[h] == k) rank[h] = __popcll(b & ((uint64_t(1) << lane) - 1));
        }
    }
    __syncthreads();
    if (threadIdx.x < kSellKeys) {   // per key: rows of the earlier chunks
        const int k = threadIdx.x;
        int acc = 0;
        for (int c = 0; c < kSellChunks; ++c) {
            const int t = cnt[c][k];
            cnt[c][k] = acc;
            acc += t;
        }
        tot[k] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {   // longest key first
        int acc = 0;
        for (int k = kSellKeys - 1; k >= 0; --k) {
            start[k] = acc;
            acc += tot[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < kPer; ++h) {
        if (!valid[h]) continue;
        const int chunk = wave + waves * h;
        const int pos = start[key[h]] + cnt[chunk][key[h]] + rank[h];
        perm[w0 + pos] = (int32_t)(w0 + (int64_t)chunk * kSellC + lane);
        len[w0 + pos] = ln[h];
    }
}

// width[s] = the longest row of slice s (one wavefront per slice)
__global__ __launch_bounds__(256) void k_sell_width(const uint32_t *len, int64_t rows, int64_t slices, int32_t *width) {
    const int64_t s = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (s >= slices) return;
    const int64_t i = s * kSellC + threadIdx.x % kWave;
    uint32_t v = i < rows ? len[i] : 0u;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)v, off);
        v = o > v ? o : v;
    }
    if (threadIdx.x % kWave == 0) width[s] = (int32_t)v;
}

// the sliced arrays from a one-lane-per-row ELL matrix (padding: column 0,
// value 0, never added)
template <typename VT>
__global__ __launch_bounds__(256) void k_sell_fill(const int32_t *Aj, const VT *Ax, int64_t stride, int64_t rows,
                                                   const int32_t *perm, const uint32_t *len, const int32_t *width,
                                                   const int64_t *base, int32_t *sAj, VT *sAx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = i / kSellC;
    const int lane = (int)(i % kSellC);
    if (s * kSellC >= rows) return;
    const bool valid = i < rows;
    const int64_t row = valid ? perm[i] : 0;
    const uint32_t l = valid ? len[i] : 0u;
    const int w = width[s];
    const int64_t b = base[s];
    for (int k = 0; k < w; ++k) {
        const bool have = (uint32_t)k < l;
        sAj[b + (int64_t)k * kSellC + lane] = have ? Aj[row + (int64_t)k * stride] : 0;
        sAx[b + (int64_t)k * kSellC + lane] = have ? Ax[row + (int64_t)k * stride] : VT(0);
    }
}

// y = A x: one wavefront per slice, four slices per work-group, work-groups
// of neighbouring slices on one XCD; the row's terms in stored order, the
// index and value loads of U slots issued before their sums.
template <typename VT>
__global__ __launch_bounds__(256) void k_sell_spmv(const int32_t *__restrict__ sAj, const VT *__restrict__ sAx,
                                                   const int32_t *__restrict__ perm, const uint32_t *__restrict__ len,
                                                   const int32_t *__restrict__ width, const int64_t *__restrict__ base,
                                                   int64_t rows, int64_t slices, const VT *__restrict__ x,
                                                   VT *__restrict__ y) {
    constexpr int U = sizeof(VT) == 8 ? 4 : 8;
    const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
    const int64_t s = lb * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (s >= slices) return;
    const int lane = threadIdx.x % kWave;
    const int64_t i = s * kSellC + lane;
    const bool valid = i < rows;
    const uint32_t l = valid ? len[i] : 0u;
    const int w = width[s];
    const int64_t b = base[s] + lane;
    VT acc = VT(0);
    int k = 0;
    for (; k + U <= w; k += U) {
        int32_t j[U];
        VT v[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            j[u] = __builtin_nontemporal_load(sAj + b + (int64_t)(k + u) * kSellC);
            v[u] = __builtin_nontemporal_load(sAx + b + (int64_t)(k + u) * kSellC);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[j[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) acc = ((uint32_t)(k + u) < l) ? acc + v[u] * xv[u] : acc;
    }
    for (; k < w; ++k) {
        const int32_t j = __builtin_nontemporal_load(sAj + b + (int64_t)k * kSellC);
        const VT v = __builtin_nontemporal_load(sAx + b + (int64_t)k * kSellC);
        acc = ((uint32_t)k < l) ? acc + v * x[j] : acc;
    }
    if (valid) y[perm[i]] = acc;
}

// Column range [lo, hi] of the stored entries (k < len) of sort window
// blockIdx.x (its 32 slices); a window without entries gets lo > hi.
__global__ __launch_bounds__(1024) void k_sell_ranges(const int32_t *__restrict__ sAj, const uint32_t *__restrict__ len,
                                                      const int32_t *__restrict__ width,
                                                      const int64_t *__restrict__ base, int64_t rows, int64_t slices,
                                                      int32_t *ranges) {
    __shared__ int32_t rl[1024 / kWave], rh[1024 / kWave];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    constexpr int kSlicesPerWave = kSellChunks / (1024 / kWave);
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int h = 0; h < kSlicesPerWave; ++h) {
        const int64_t sl = (int64_t)blockIdx.x * kSellChunks + wave * kSlicesPerWave + h;
        if (sl >= slices) break;
        const int64_t i = sl * kSellC + lane;
        const uint32_t l = i < rows ? len[i] : 0u;
        const int64_t b = base[sl] + lane;
        for (uint32_t k = 0; k < l; ++k) {
            const int32_t j = sAj[b + (int64_t)k * kSellC];
            lo = j < lo ? j : lo;
            hi = j > hi ? j : hi;
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int32_t ol = __shfl_xor(lo, off), oh = __shfl_xor(hi, off);
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
    }
    if (lane == 0) {
        rl[wave] = lo;
        rh[wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 1024 / kWave; ++w) {
            lo = rl[w] < lo ? rl[w] : lo;
            hi = rh[w] > hi ? rh[w] : hi;
        }
        ranges[2 * blockIdx.x] = lo;
        ranges[2 * blockIdx.x + 1] = hi;
    }
}

// y = A x with x staged in LDS: one 1024-thread work-group per sort window
// (32 slices, two per wavefront).  The window's column range of x is copied
// into LDS once (coalesced), then each slice streams its index and value
// arrays and gathers x from LDS.  The terms and their order are k_sell_spmv's,
// so y has the same bits; a window wider than `cap` entries, or a row with an
// entry outside the recorded range, sums that row from global x instead.
template <typename VT>
__global__ __launch_bounds__(1024) void k_sell_spmv_staged(
    const int32_t *__restrict__ sAj, const VT *__restrict__ sAx, const int32_t *__restrict__ perm,
    const uint32_t *__restrict__ len, const int32_t *__restrict__ width, const int64_t *__restrict__ base,
    int64_t rows, int64_t slices, const VT *__restrict__ x, VT *__restrict__ y, const int32_t *__restrict__ ranges,
    int64_t cap) {
    extern __shared__ unsigned char sell_stage_raw[];
    VT *win = reinterpret_cast<VT *>(sell_stage_raw);
    constexpr int U = sizeof(VT) == 8 ? 4 : 8;
    constexpr int kSlicesPerWave = kSellChunks / (1024 / kWave);
    const int64_t wdw = xcd_block(blockIdx.x, gridDim.x);   // neighbouring windows on one XCD
    const int32_t lo = ranges[2 * wdw], hi = ranges[2 * wdw + 1];
    const int64_t wn = (int64_t)hi - (int64_t)lo + 1;
    const bool staged = hi >= lo && wn <= cap;
    if (staged)
        for (int64_t i = threadIdx.x; i < wn; i += blockDim.x) win[i] = x[lo + i];
    __syncthreads();
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    for (int h = 0; h < kSlicesPerWave; ++h) {
        const int64_t sl = wdw * kSellChunks + wave * kSlicesPerWave + h;
        if (sl >= slices) break;
        const int64_t i = sl * kSellC + lane;
        const bool valid = i < rows;
        const uint32_t l = valid ? len[i] : 0u;
        const int w = width[sl];
        const int64_t b = base[sl] + lane;
        VT acc = VT(0);
        bool outside = !staged;
        if (staged) {
            int k = 0;
            for (; k + U <= w; k += U) {
                int32_t j[U];
                VT v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    j[u] = __builtin_nontemporal_load(sAj + b + (int64_t)(k + u) * kSellC);
                    v[u] = __builtin_nontemporal_load(sAx + b + (int64_t)(k + u) * kSellC);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t o = (uint32_t)(j[u] - lo);
                    const bool in = o < (uint32_t)wn, use = (uint32_t)(k + u) < l;
                    outside |= use && !in;
                    acc = use ? acc + v[u] * win[in ? o : 0u] : acc;
                }
            }
            for (; k < w; ++k) {
                const int32_t j = __builtin_nontemporal_load(sAj + b + (int64_t)k * kSellC);
                const VT v = __builtin_nontemporal_load(sAx + b + (int64_t)k * kSellC);
                const uint32_t o = (uint32_t)(j - lo);
                const bool in = o < (uint32_t)wn, use = (uint32_t)k < l;
                outside |= use && !in;
                acc = use ? acc + v * win[in ? o : 0u] : acc;
            }
        }
        if (outside) {   // the row from global x, the same terms in the same order
            acc = VT(0);
            for (uint32_t k = 0; k < l; ++k) acc = acc + sAx[b + (int64_t)k * kSellC] * x[sAj[b + (int64_t)k * kSellC]];
        }
        if (valid) y[perm[i]] = acc;
    }
}

}  // namespace

int sell_window() { return kSellWindow; }
int sell_slice_rows() { return kSellC; }

hipError_t sell_sort(const uint32_t *Arl, int64_t rows, int32_t *perm, uint32_t *len, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    const int64_t windows = (rows + kSellWindow - 1) / kSellWindow;
    if (windows > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sell_sort, dim3((unsigned)windows), dim3(1024), 0, s, Arl, rows, perm, len);
    return hipGetLastError();
}

hipError_t sell_widths(const uint32_t *len, int64_t rows, int64_t slices, int32_t *width, hipStream_t s) {
    if (slices <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sell_width, dim3((unsigned)((slices + 3) / 4)), dim3(256), 0, s, len, rows, slices, width);
    return hipGetLastError();
}

hipError_t sell_fill(ValueType vt, const int32_t *Aj, const void *Ax, int64_t stride, int64_t rows,
                     const int32_t *perm, const uint32_t *len, const int32_t *width, const int64_t *base,
                     int32_t *sAj, void *sAx, hipStream_t s) {
    const int64_t slices = (rows + kSellC - 1) / kSellC;
    if (slices <= 0) return hipSuccess;
    const dim3 g((unsigned)((slices * kSellC + 255) / 256)), b(256);
    if (vt == VT_F32)
        hipLaunchKernelGGL(k_sell_fill<float>, g, b, 0, s, Aj, (const float *)Ax, stride, rows, perm, len, width, base,
                           sAj, (float *)sAx);
    else if (vt == VT_F64)
        hipLaunchKernelGGL(k_sell_fill<double>, g, b, 0, s, Aj, (const double *)Ax, stride, rows, perm, len, width,
                           base, sAj, (double *)sAx);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t sell_ranges(const int32_t *sAj, const uint32_t *len, const int32_t *width, const int64_t *base,
                       int64_t rows, int32_t *ranges, hipStream_t s) {
    const int64_t slices = (rows + kSellC - 1) / kSellC;
    const int64_t windows = (rows + kSellWindow - 1) / kSellWindow;
    if (windows <= 0) return hipSuccess;
    if (windows > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sell_ranges, dim3((unsigned)windows), dim3(1024), 0, s, sAj, len, width, base, rows, slices,
                       ranges);
    return hipGetLastError();
}

hipError_t sell_spmv_staged(ValueType vt, const int32_t *sAj, const void *sAx, const int32_t *perm,
                            const uint32_t *len, const int32_t *width, const int64_t *base, int64_t rows,
                            const void *x, void *y, const int32_t *ranges, int64_t window, hipStream_t s) {
    const int64_t slices = (rows + kSellC - 1) / kSellC;
    const int64_t windows = (rows + kSellWindow - 1) / kSellWindow;
    if (windows <= 0) return hipSuccess;
    const size_t vs = vt == VT_F64 ? 8 : 4;
    if (window < 1 || (size_t)window * vs > kMaxStageBytes || windows > 0x7fffffff) return hipErrorInvalidValue;
    const size_t lds = (size_t)window * vs;
    if (vt == VT_F32)
        hipLaunchKernelGGL(k_sell_spmv_staged<float>, dim3((unsigned)windows), dim3(1024), lds, s, sAj,
                           (const float *)sAx, perm, len, width, base, rows, slices, (const float *)x, (float *)y,
                           ranges, window);
    else if (vt == VT_F64)
        hipLaunchKernelGGL(k_sell_spmv_staged<double>, dim3((unsigned)windows), dim3(1024), lds, s, sAj,
                           (const double *)sAx, perm, len, width, base, rows, slices, (const double *)x, (double *)y,
                           ranges, window);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t sell_spmv(ValueType vt, const int32_t *sAj, const void *sAx, const int32_t *perm, const uint32_t *len,
                     const int32_t *width, const int64_t *base, int64_t rows, const void *x, void *y, hipStream_t s) {
    const int64_t slices = (rows + kSellC - 1) / kSellC;
    if (slices <= 0) return hipSuccess;
    const int64_t blocks = (slices + 3) / 4;
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;
    if (vt == VT_F32)
        hipLaunchKernelGGL(k_sell_spmv<float>, dim3((unsigned)blocks), dim3(256), 0, s, sAj, (const float *)sAx, perm,
                           len, width, base, rows, slices, (const float *)x, (float *)y);
    else if (vt == VT_F64)
        hipLaunchKernelGGL(k_sell_spmv<double>, dim3((unsigned)blocks), dim3(256), 0, s, sAj, (const double *)sAx,
                           perm, len, width, base, rows, slices, (const double *)x, (double *)y);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace hipx
}  // namespace honei
