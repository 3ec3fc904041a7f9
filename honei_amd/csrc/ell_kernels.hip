// ell_kernels.hip -- ELL row kernels for gfx950: SpMV, defect, fused
// defect+norm and fused Jacobi sweeps.
//
// One row is served by T consecutive lanes of a 64-wide wavefront (T a power
// of two, T <= 64, matching the matrix's `threads` layout).  Lane t of row r
// walks slots r*T + t + k*stride for k < Arl[r]; consecutive lanes therefore
// read consecutive addresses of each column-major ELL stream, which the
// hardware coalesces into full 64-byte-line requests.  Ax/Aj are streamed with
// non-temporal loads so that the gathered vector keeps L2/MALL.  Partial sums
// of the T lanes are combined with xor-butterflies, which for lane 0 adds the
// same pairs in the same order as the reference's shared-memory tree
// (backends/cuda/product.cu:176-198).  Multiply and add stay separate
// roundings (no FMA contraction) so that for T = 1 the result is
// bit-identical to Product<tags::CPU> (la/product.hh:148-206).

#include <algorithm>
#include <type_traits>

#include "backend.hh"
#include "ell_device.hh"

#pragma clang fp contract(off)

namespace honei {
namespace hipx {
namespace {

struct KArgs {
    const int32_t *Aj;
    const void *Ax;
    const uint32_t *Arl;
    int64_t stride;
    void *out;
    const void *x;
    const void *rhs;
    const void *dinv;
    void *acc;
    double *norm2;
    int64_t row_start, row_end;
    int64_t nlogical;
    ReduceSlots slots;
    uint32_t aj_bytes, ax_bytes;   // buffer extents of the fixed-width form
    const int *gate;               // non-null: run only while *gate != 0
};

// One block serves `per_block` consecutive logical blocks of blockDim/T rows
// (per_block = 1 unless the launch was capped, as it is for launches that
// carry a reduction: few blocks keep the arrival counter cold).  W: the
// stored width is q1_width(T) (the Q1 operators among others) and the row
// sum takes the fixed-width form through buffer loads.  IW: 32-bit words per
// stored index in Aj and Arl (1: int32 arrays; 2: the reference's 64-bit
// arrays, low word read).
template <typename VT, int T, int MODE, bool NORM, bool ACC, bool W, int IW = 1>
__global__ __launch_bounds__(1024) void ell_row_kernel(KArgs a) {
    typedef typename VecOf<VT>::type V;
    const VT *__restrict__ Ax = (const VT *)a.Ax;
    const int32_t *__restrict__ Aj = a.Aj;
    const V *__restrict__ x = (const V *)a.x;
    const V *__restrict__ rhs = (const V *)a.rhs;
    const V *__restrict__ dinv = (const V *)a.dinv;
    if (a.gate && !block_uniform_flag(a.gate)) return;

    const int64_t p = xcd_block(blockIdx.x, gridDim.x);
    const int64_t per = a.nlogical / gridDim.x, rem = a.nlogical % gridDim.x;
    const int64_t first = p * per + (p < rem ? p : rem);
    const int64_t last = first + per + (p < rem ? 1 : 0);

    double nval = 0.0;
    for (int64_t lb = first; lb < last; ++lb) {
        const int64_t gid = lb * (int64_t)blockDim.x + threadIdx.x;
        const int lane = (int)(gid & (T - 1));
        const int64_t row = a.row_start + gid / T;
        const bool active = row < a.row_end;

        V s = V(0);
        if (active) {
            const uint32_t len = a.Arl[row * IW];
            if constexpr (W) {
                const auto rAj = buffer_rsrc(Aj, a.aj_bytes), rAx = buffer_rsrc(Ax, a.ax_bytes);
                if constexpr (MODE == ELL_JACOBI0)
                    s = ell_lane_sum_buf<VT, T, q1_width(T), true, IW>(rAj, rAx, row, lane, len, a.stride, rhs, dinv);
                else
                    s = ell_lane_sum_buf<VT, T, q1_width(T), false, IW>(rAj, rAx, row, lane, len, a.stride, x,
                                                                         nullptr);
            } else {
                if constexpr (MODE == ELL_JACOBI0)
                    s = ell_lane_sum<VT, T, true, false, IW>(Aj, Ax, row, lane, len, a.stride, rhs, dinv);
                else
                    s = ell_lane_sum<VT, T, false, false, IW>(Aj, Ax, row, lane, len, a.stride, x);
            }
        }
        s = lane_group_sum<T>(s);

        if (active && lane == 0) {
            if constexpr (MODE == ELL_PRODUCT) {
                ((V *)a.out)[row - a.row_start] = s;
            } else if constexpr (MODE == ELL_DEFECT) {
                const V r = rhs[row] - s;
                ((V *)a.out)[row] = r;
                if constexpr (NORM) nval += (double)r * (double)r;
            } else {
                const V xin = (MODE == ELL_JACOBI0) ? dinv[row] * rhs[row] : x[row];
                const V t = rhs[row] - s;
                const V o = xin + t * dinv[row];
                ((V *)a.out)[row] = o;
                if constexpr (ACC) {
                    V *acc = (V *)a.acc;
                    acc[row] = acc[row] + o;
                }
            }
        }
    }
    if constexpr (NORM) {
        const double v = block_sum(nval);
        if (threadIdx.x == 0) a.slots.partials[blockIdx.x] = v;
    }
}

__global__ void k_finish_norm(const double *partials, int64_t nparts, double *out, const int *gate) {
    if (gate && !block_uniform_flag(gate)) return;
    const double t = sum_partials(partials, nparts, 1, 0);
    if (threadIdx.x == 0) *out = t;
}

typedef void (*KernelFn)(KArgs);

template <typename VT, int T, bool W>
KernelFn pick_mode_w(EllMode mode, bool norm, bool acc, int iw) {
    if (iw == 2) {   // the reference's 64-bit index arrays: Product and Defect
        if constexpr (!std::is_same<VT, bf16>::value) {
            if (mode == ELL_PRODUCT) return ell_row_kernel<VT, T, ELL_PRODUCT, false, false, W, 2>;
            if (mode == ELL_DEFECT && !norm) return ell_row_kernel<VT, T, ELL_DEFECT, false, false, W, 2>;
        }
        return nullptr;
    }
    if (iw != 1) return nullptr;
    switch (mode) {
        case ELL_PRODUCT: return ell_row_kernel<VT, T, ELL_PRODUCT, false, false, W>;
        case ELL_DEFECT:
            return norm ? ell_row_kernel<VT, T, ELL_DEFECT, true, false, W>
                        : ell_row_kernel<VT, T, ELL_DEFECT, false, false, W>;
        case ELL_JACOBI:
            return acc ? ell_row_kernel<VT, T, ELL_JACOBI, false, true, W>
                       : ell_row_kernel<VT, T, ELL_JACOBI, false, false, W>;
        case ELL_JACOBI0:
            return acc ? ell_row_kernel<VT, T, ELL_JACOBI0, false, true, W>
                       : ell_row_kernel<VT, T, ELL_JACOBI0, false, false, W>;
    }
    return nullptr;
}

template <typename VT, int T>
KernelFn pick_mode(EllMode mode, bool norm, bool acc, bool wide, int iw) {
    if constexpr (T <= 8)
        if (wide) return pick_mode_w<VT, T, true>(mode, norm, acc, iw);
    return pick_mode_w<VT, T, false>(mode, norm, acc, iw);
}

template <typename VT>
KernelFn pick_threads(int T, EllMode mode, bool norm, bool acc, bool wide, int iw) {
    switch (T) {
        case 1: return pick_mode<VT, 1>(mode, norm, acc, wide, iw);
        case 2: return pick_mode<VT, 2>(mode, norm, acc, wide, iw);
        case 4: return pick_mode<VT, 4>(mode, norm, acc, wide, iw);
        case 8: return pick_mode<VT, 8>(mode, norm, acc, wide, iw);
        case 16: return pick_mode<VT, 16>(mode, norm, acc, wide, iw);
        case 32: return pick_mode<VT, 32>(mode, norm, acc, wide, iw);
        case 64: return pick_mode<VT, 64>(mode, norm, acc, wide, iw);
    }
    return nullptr;
}

// ---- x staged in LDS ----------------------------------------------------------
// Column window [min, max] of the stored entries k < Arl[r] of the rows of
// block b (rows_per_block rows), column 0 left out: padding slots inside a
// row's last T-wide chunk hold column 0, and the staged kernel keeps x[0] in
// a register.  A block without other columns gets min > max.
__global__ __launch_bounds__(256) void k_block_ranges(const int32_t *Aj, const uint32_t *Arl, int64_t rows,
                                                      int64_t stride, int T, int64_t rpb, int32_t *ranges) {
    const int64_t b = blockIdx.x;
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    const int64_t r0 = b * rpb, slots = rpb * T;
    for (int64_t i = threadIdx.x; i < slots; i += blockDim.x) {
        const int64_t row = r0 + i / T;
        if (row >= rows) break;
        const int64_t lane = i % T;
        const uint32_t len = Arl[row];
        for (uint32_t k = 0; k < len; ++k) {
            const int32_t j = Aj[row * T + lane + (int64_t)k * stride];
            if (j == 0) continue;
            lo = j < lo ? j : lo;
            hi = j > hi ? j : hi;
        }
    }
    __shared__ int32_t rl[256 / kWave], rh[256 / kWave];
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int32_t ol = __shfl_xor(lo, off), oh = __shfl_xor(hi, off);
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
    }
    if (threadIdx.x % kWave == 0) {
        rl[threadIdx.x / kWave] = lo;
        rh[threadIdx.x / kWave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / kWave); ++w) {
            lo = rl[w] < lo ? rl[w] : lo;
            hi = rh[w] > hi ? rh[w] : hi;
        }
        ranges[2 * b] = lo;
        ranges[2 * b + 1] = hi;
    }
}

// y = A x, one block of `chunks` x blockDim/T rows per work-group: the
// block's x window is copied into LDS (coalesced) once and the row sums of
// its chunks gather from it; a window wider than `cap` elements is gathered
// from global memory.  The row sum is ell_lane_sum's, term for term, so y
// is bit-identical to ELL_PRODUCT.
template <typename VT, int T>
__global__ __launch_bounds__(1024) void k_product_staged(const int32_t *__restrict__ Aj, const VT *__restrict__ Ax,
                                                         const uint32_t *__restrict__ Arl, int64_t stride, int64_t rows,
                                                         const typename VecOf<VT>::type *__restrict__ x,
                                                         typename VecOf<VT>::type *__restrict__ y,
                                                         const int32_t *__restrict__ ranges, int64_t cap, int chunks) {
    typedef typename VecOf<VT>::type V;
    extern __shared__ unsigned char stage_raw[];
    V *win = reinterpret_cast<V *>(stage_raw);
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);   // neighbouring windows on one XCD
    const int32_t lo = ranges[2 * b], hi = ranges[2 * b + 1];
    const int64_t width = (int64_t)hi - (int64_t)lo + 1;
    const bool staged = hi >= lo && width <= cap;
    if (staged)
        for (int64_t i = threadIdx.x; i < width; i += blockDim.x) win[i] = x[lo + i];
    __syncthreads();
    const V x0 = x[0];
    const int lane = (int)(threadIdx.x & (T - 1));
    const int64_t rows_per_chunk = blockDim.x / T;
    for (int c = 0; c < chunks; ++c) {
        const int64_t row = (b * chunks + c) * rows_per_chunk + threadIdx.x / T;
        V s = V(0);
        if (row < rows) {
            const uint32_t len = Arl[row];
            int64_t slot = row * T + lane;
            bool outside = !staged;
            if (staged) {
                // column 0 (padding) reads x0; any other column outside the
                // window (ranges not made for this block) sends the lane to
                // the global path.  U entries at a time (8 for 4-byte, 4 for
                // 8-byte vectors): their index and value loads are all issued
                // before the first use.
                constexpr int U = sizeof(V) == 8 ? 4 : 8;
                auto term = [&](int32_t j, V av) {
                    const uint32_t o = (uint32_t)(j - lo);
                    const bool in = o < (uint32_t)width;
                    outside |= !in && j != 0;
                    s += av * (in ? win[o] : x0);
                };
                uint32_t k = 0;
                for (; k + U <= len; k += U) {
                    int32_t j[U];
                    V av[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) j[u] = __builtin_nontemporal_load(Aj + slot + u * stride);
#pragma unroll
                    for (int u = 0; u < U; ++u) av[u] = load_value_nt(Ax + slot + u * stride);
#pragma unroll
                    for (int u = 0; u < U; ++u) term(j[u], av[u]);
                    slot += U * stride;
                }
                for (; k < len; ++k) {
                    term(__builtin_nontemporal_load(Aj + slot), load_value_nt(Ax + slot));
                    slot += stride;
                }
            }
            if (outside) s = ell_lane_sum<VT, T>(Aj, Ax, row, lane, len, stride, x);
        }
        s = lane_group_sum<T>(s);
        if (row < rows && lane == 0) y[row] = s;
    }
}

typedef void (*StagedFn)(const int32_t *, const void *, const uint32_t *, int64_t, int64_t, const void *, void *,
                         const int32_t *, int64_t, int);

template <typename VT>
StagedFn pick_staged(int T) {
    switch (T) {
        case 1: return (StagedFn)k_product_staged<VT, 1>;
        case 2: return (StagedFn)k_product_staged<VT, 2>;
        case 4: return (StagedFn)k_product_staged<VT, 4>;
        case 8: return (StagedFn)k_product_staged<VT, 8>;
        case 16: return (StagedFn)k_product_staged<VT, 16>;
        case 32: return (StagedFn)k_product_staged<VT, 32>;
        case 64: return (StagedFn)k_product_staged<VT, 64>;
    }
    return nullptr;
}

}  // namespace

hipError_t ell_block_ranges(const EllView &A, int64_t rows_per_block, int32_t *ranges, hipStream_t s) {
    const int T = A.threads;
    if (A.index_words != 1) return hipErrorInvalidValue;
    if (rows_per_block < 1 || T < 1 || T > 64 || (T & (T - 1)) || !ranges) return hipErrorInvalidValue;
    const int64_t nb = (A.rows + rows_per_block - 1) / rows_per_block;
    if (nb <= 0) return hipSuccess;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_block_ranges, dim3((unsigned)nb), dim3(256), 0, s, A.Aj, A.Arl, A.rows, A.stride, T,
                       rows_per_block, ranges);
    return hipGetLastError();
}

hipError_t ell_product_staged(const EllView &A, void *y, const void *x, const int32_t *ranges, int block, int chunks,
                              int64_t window, hipStream_t s) {
    const int T = A.threads;
    StagedFn fn = nullptr;
    switch (A.vtype) {
        case VT_F32: fn = pick_staged<float>(T); break;
        case VT_F64: fn = pick_staged<double>(T); break;
        case VT_BF16: fn = pick_staged<bf16>(T); break;
    }
    if (!fn || block % kWave || block > 1024 || block < T || !ranges || chunks < 1 || A.index_words != 1)
        return hipErrorInvalidValue;
    const int64_t rpb = (int64_t)(block / T) * chunks;
    const int64_t nb = (A.rows + rpb - 1) / rpb;
    if (nb <= 0) return hipSuccess;
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    const size_t vs = A.vtype == VT_F64 ? 8 : 4;
    if (window < 1 || (size_t)window * vs > kMaxStageBytes) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3((unsigned)nb), dim3(block), (size_t)window * vs, s, A.Aj, A.Ax, A.Arl, A.stride, A.rows,
                       x, y, ranges, window, chunks);
    return hipGetLastError();
}

hipError_t ell_apply(const EllView &A, const EllArgs &e, const ReduceSlots &slots, hipStream_t s) {
    const int T = A.threads;
    const bool norm = (e.mode == ELL_DEFECT) && e.norm2 != nullptr;
    const bool acc = (e.mode == ELL_JACOBI || e.mode == ELL_JACOBI0) && e.acc != nullptr;
    // fixed-width form: stored width q1_width(T) and every byte offset within 31 bits
    const int iw = A.index_words;
    const int64_t vsz = A.vtype == VT_F64 ? 8 : (A.vtype == VT_BF16 ? 2 : 4);
    const int64_t slots_total = A.ncpr * A.stride;
    const bool wide = T <= 8 && A.ncpr == q1_width(T) && slots_total * 8 < (int64_t(1) << 31);
    KernelFn fn = nullptr;
    switch (A.vtype) {
        case VT_F32: fn = pick_threads<float>(T, e.mode, norm, acc, wide, iw); break;
        case VT_F64: fn = pick_threads<double>(T, e.mode, norm, acc, wide, iw); break;
        case VT_BF16: fn = pick_threads<bf16>(T, e.mode, norm, acc, wide, iw); break;
    }
    if (!fn) return hipErrorInvalidValue;
    const int64_t nrows = e.row_end - e.row_start;
    int block = e.block > 0 ? e.block : 256;
    if (block % kWave != 0 || block > 1024 || block < T) return hipErrorInvalidValue;
    if (nrows <= 0) {
        if (norm) return fill_bytes_zero(e.norm2, sizeof(double), s);
        return hipSuccess;
    }
    const int64_t nlogical = (nrows * T + block - 1) / block;
    if (nlogical > 0x7fffffff) return hipErrorInvalidValue;
    // launches that reduce keep the block count within the partial slots
    const int64_t grid = (norm && nlogical > slots.capacity) ? slots.capacity : nlogical;
    KArgs k;
    k.Aj = A.Aj;
    k.Ax = A.Ax;
    k.Arl = A.Arl;
    k.stride = A.stride;
    k.out = e.out;
    k.x = e.x;
    k.rhs = e.rhs;
    k.dinv = e.dinv;
    k.acc = e.acc;
    k.norm2 = e.norm2;
    k.row_start = e.row_start;
    k.row_end = e.row_end;
    k.nlogical = nlogical;
    k.slots = slots;
    k.aj_bytes = (uint32_t)(slots_total * 4 * iw);
    k.gate = e.gate;
    k.ax_bytes = (uint32_t)(slots_total * vsz);
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(block), 0, s, k);
    if (norm)
        hipLaunchKernelGGL(k_finish_norm, dim3(1), dim3(1024), 0, s, (const double *)slots.partials, grid, e.norm2,
                           e.gate);
    return hipGetLastError();
}

}  // namespace hipx
}  // namespace honei
