// generators.hip -- synthetic workloads generated in HBM.
//
// Random ELL matrix of the SpMV benchmark configuration, T = 1: row r gets
// L_r ~ U{1..max_len} distinct columns uniform in [r-band, r+band] clipped
// to [0, n), sorted ascending, values N(0,1) (Box-Muller), laid out exactly
// as SparseMatrixELL(SparseMatrix) stores it (la/sparse_matrix_ell-impl.hh:
// 75-122): slot r + k*stride, zero padding, stride 32-aligned by the caller.
// Draws come from a SplitMix64 stream per row keyed by (seed, r), so the
// matrix depends only on (n, max_len, band, seed).
//
// Fields for the Poisson workloads: uniform right-hand sides and lognormal
// cell coefficients, element i drawn from position i of one SplitMix64
// sequence, so any element can be generated independently.

#include "q1_device.hh"
#include "solvers.hh"

namespace honei {
namespace hipx {
namespace {

__device__ inline uint64_t splitmix_next(uint64_t &state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ inline double u01(uint64_t &st) { return (double)(splitmix_next(st) >> 11) * (1.0 / 9007199254740992.0); }

template <typename VT, int MAXW>
__global__ void k_random_ell(int64_t n, int max_len, int64_t band, uint64_t seed, int32_t *Aj, VT *Ax, uint32_t *Arl,
                             int64_t stride, int T) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nchunks = (max_len + T - 1) / T;
    if (r * T >= stride) return;
    if (r >= n) {   // padding slots beyond the last row
        for (int64_t k = 0; k < nchunks; ++k)
            for (int t = 0; t < T && r * T + t < stride; ++t) {
                Aj[r * T + t + k * stride] = 0;
                Ax[r * T + t + k * stride] = VT(0);
            }
        return;
    }
    uint64_t st = seed ^ ((uint64_t)r * 0xD1B54A32D192ED03ull);
    (void)splitmix_next(st);
    const int L = 1 + (int)(splitmix_next(st) % (uint64_t)max_len);
    const int64_t lo = r - band < 0 ? 0 : r - band;
    const int64_t hi = r + band > n - 1 ? n - 1 : r + band;
    const int64_t width = hi - lo + 1;
    int32_t cols[MAXW];
    int have = 0;
    for (int tries = 0; have < L && tries < 64 * MAXW; ++tries) {
        const int32_t c = (int32_t)(lo + (int64_t)(u01(st) * (double)width));
        // insertion into the sorted prefix unless already present
        int pos = have;
        bool dup = false;
        while (pos > 0 && cols[pos - 1] >= c) {
            if (cols[pos - 1] == c) {
                dup = true;
                break;
            }
            --pos;
        }
        if (dup) continue;
        for (int q = have; q > pos; --q) cols[q] = cols[q - 1];
        cols[pos] = c;
        ++have;
    }
    // entry t of the row goes to slot (t % T) + r*T + (t / T)*stride
    for (int t = 0; t < nchunks * T; ++t) {
        const int64_t slot = (t % T) + r * T + (int64_t)(t / T) * stride;
        if (t < have) {
            double v;
            do {
                const double u1 = u01(st), u2 = u01(st);
                v = sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
            } while (v == 0.0);
            Aj[slot] = cols[t];
            Ax[slot] = (VT)v;
        } else {
            Aj[slot] = 0;
            Ax[slot] = VT(0);
        }
    }
    Arl[r] = (uint32_t)((have + T - 1) / T);
}

// dst[i] = lo + (hi - lo) * U_i
__global__ void k_fill_uniform(double *dst, int64_t n, double lo, double hi, uint64_t seed) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = uniform_at(lo, hi, seed, (uint64_t)i);
}

// dst[i] = exp(sigma * z_i) (lognormal_at)
__global__ void k_fill_lognormal(double *dst, int64_t n, double sigma, uint64_t seed) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = lognormal_at(sigma, seed, (uint64_t)i);
}

inline int field_grid(int64_t n) {
    int64_t g = (n + 1023) / 1024;
    return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

}  // namespace

hipError_t random_ell(ValueType vt, int64_t n, int max_len, int64_t band, uint64_t seed, int32_t *Aj, void *Ax,
                      uint32_t *Arl, int64_t stride, int threads, hipStream_t s) {
    if (max_len < 1 || max_len > 64 || stride < n * threads || band < max_len || threads < 1 || threads > 64)
        return hipErrorInvalidValue;
    const int64_t slots_rows = (stride + threads - 1) / threads;
    const int grid = (int)((slots_rows + 255) / 256);
    if (vt == VT_F32)
        hipLaunchKernelGGL((k_random_ell<float, 64>), dim3(grid), dim3(256), 0, s, n, max_len, band, seed, Aj,
                           (float *)Ax, Arl, stride, threads);
    else if (vt == VT_F64)
        hipLaunchKernelGGL((k_random_ell<double, 64>), dim3(grid), dim3(256), 0, s, n, max_len, band, seed, Aj,
                           (double *)Ax, Arl, stride, threads);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t fill_uniform(double *dst, int64_t n, double lo, double hi, uint64_t seed, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_uniform, dim3(field_grid(n)), dim3(256), 0, s, dst, n, lo, hi, seed);
    return hipGetLastError();
}

hipError_t fill_lognormal(double *dst, int64_t n, double sigma, uint64_t seed, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_lognormal, dim3(field_grid(n)), dim3(256), 0, s, dst, n, sigma, seed);
    return hipGetLastError();
}

}  // namespace hipx
}  // namespace honei
