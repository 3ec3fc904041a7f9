// capi.cpp -- extern "C" entry points declared in include/honei_hip.h.
//
// Validation happens here (before any launch), mirroring the size checks of
// the reference's glue (e.g. la/product-cuda.cc:585-592); launch errors are
// always checked with hipGetLastError, without forcing a synchronisation.

#include "honei_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "backend.hh"
#include "runtime.hh"
#include "solvers.hh"

using namespace honei::hipx;

namespace {

inline hipStream_t S(void *s) { return (hipStream_t)s; }

int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

int check(hipError_t e, const char *what) {
    if (e == hipSuccess) return HONEI_HIP_SUCCESS;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(HONEI_HIP_ERROR_OUT_OF_MEMORY, std::string(what) + ": out of device memory");
    if (e == hipErrorInvalidValue) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, std::string(what) + ": invalid argument");
    return fail(HONEI_HIP_ERROR_RUNTIME, std::string(what) + ": " + hipGetErrorString(e));
}

bool valid_vtype(int vt) { return vt == HONEI_HIP_F32 || vt == HONEI_HIP_F64; }
bool pow2_threads(size_t t) { return t == 1 || t == 2 || t == 4 || t == 8 || t == 16 || t == 32 || t == 64; }

int check_ell(const void *Aj, const void *Ax, const void *Arl, size_t row_start, size_t row_end, size_t stride,
              size_t threads, size_t blocksize) {
    if (!Aj || !Ax || !Arl) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "ELL arrays must be non-null");
    if (!pow2_threads(threads)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "threads must be a power of two <= 64");
    if (row_end < row_start) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "row_end < row_start");
    if (row_end * threads > stride && row_end > 0)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "stride smaller than rows*threads");
    if (blocksize != 0 && (blocksize % 64 != 0 || blocksize > 1024 || blocksize < threads))
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "blocksize must be a multiple of 64 in [threads, 1024]");
    if (row_end > (size_t)INT32_MAX) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "rows exceed 32-bit index range");
    return HONEI_HIP_SUCCESS;
}

EllView make_view(const void *Aj, const void *Ax, const void *Arl, size_t rows, size_t ncpr, size_t stride,
                  size_t threads, ValueType vt, int index_words = 1) {
    EllView A;
    A.rows = (int64_t)rows;
    A.cols = 0;
    A.stride = (int64_t)stride;
    A.ncpr = (int64_t)ncpr;
    A.threads = (int)threads;
    A.Aj = (const int32_t *)Aj;
    A.Ax = Ax;
    A.Arl = (const uint32_t *)Arl;
    A.vtype = vt;
    A.index_words = index_words;
    return A;
}

int ell_call(ValueType vt, EllArgs a, const void *Aj, const void *Ax, const void *Arl, size_t ncpr, size_t stride,
             size_t threads, size_t blocksize, void *stream, const char *what, int index_words = 1) {
    int st = check_ell(Aj, Ax, Arl, (size_t)a.row_start, (size_t)a.row_end, stride, threads, blocksize);
    if (st) return st;
    Workspace *w = workspace_for(S(stream));
    if (!w) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    a.block = (int)blocksize;
    EllView A = make_view(Aj, Ax, Arl, (size_t)a.row_end, ncpr, stride, threads, vt, index_words);
    return check(ell_apply(A, a, w->slots, S(stream)), what);
}

// Reduction into workspace scalar 0, then read back on the host.
int reduce_to_host(double *out, void *stream, hipError_t (*launch)(double *, const ReduceSlots &, void *),
                   void *ctx) {
    Workspace *w = workspace_for(S(stream));
    if (!w) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    int st = check(launch(w->scalars, w->slots, ctx), "reduction");
    if (st) return st;
    double v = 0;
    st = check(hipMemcpyAsync(&v, w->scalars, sizeof(double), hipMemcpyDeviceToHost, S(stream)), "reduction readback");
    if (st) return st;
    st = check(hipStreamSynchronize(S(stream)), "reduction sync");
    if (st) return st;
    *out = v;
    return HONEI_HIP_SUCCESS;
}

struct RedCtx {
    const void *x, *y;
    int64_t n;
    ValueType vt;
    hipStream_t s;
};
hipError_t launch_dot(double *out, const ReduceSlots &sl, void *c) {
    RedCtx *r = (RedCtx *)c;
    return dot(r->x, r->y, r->n, r->vt, out, sl, r->s);
}
hipError_t launch_norm(double *out, const ReduceSlots &sl, void *c) {
    RedCtx *r = (RedCtx *)c;
    return norm2(r->x, r->n, r->vt, out, sl, r->s);
}

uint16_t f32_to_bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);  // NaN
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

}  // namespace

extern "C" {

const char *honei_hip_last_error(void) { return last_error(); }
const char *honei_hip_version(void) { return "honei-hip 0.1 (gfx950)"; }

int honei_hip_device_count(int *count) {
    if (!count) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(HONEI_HIP_ERROR_NO_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *count = n;
    return n > 0 ? HONEI_HIP_SUCCESS : fail(HONEI_HIP_ERROR_NO_DEVICE, "no HIP device");
}

int honei_hip_init(int device) {
    int n = 0;
    int st = honei_hip_device_count(&n);
    if (st) return st;
    if (device < 0 || device >= n) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "device index out of range");
    st = check(hipSetDevice(device), "hipSetDevice");
    if (st) return st;
    if (!workspace_for(nullptr)) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    return noalloc_setup(nullptr);
}

int honei_hip_shutdown(void) {
    (void)hipDeviceSynchronize();
    release_workspaces();
    return HONEI_HIP_SUCCESS;
}

int honei_hip_host_alloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "ptr is null");
    *ptr = nullptr;
    return check(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
}
int honei_hip_host_free(void *ptr) { return ptr ? check(hipHostFree(ptr), "hipHostFree") : HONEI_HIP_SUCCESS; }

int honei_hip_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "ptr is null");
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    return check(hipMalloc(ptr, bytes), "hipMalloc");
}
int honei_hip_free(void *ptr) { return ptr ? check(hipFree(ptr), "hipFree") : HONEI_HIP_SUCCESS; }

int honei_hip_upload(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return HONEI_HIP_SUCCESS;
    if (!dst || !src) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    int st = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)), "upload");
    if (st) return st;
    return check(hipStreamSynchronize(S(stream)), "upload sync");
}
int honei_hip_download(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return HONEI_HIP_SUCCESS;
    if (!dst || !src) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    int st = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)), "download");
    if (st) return st;
    return check(hipStreamSynchronize(S(stream)), "download sync");
}
int honei_hip_copy(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return HONEI_HIP_SUCCESS;
    if (!dst || !src) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)), "copy");
}
int honei_hip_fill_zero(void *dst, size_t bytes, void *stream) {
    if (bytes && !dst) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    return check(fill_bytes_zero(dst, bytes, S(stream)), "fill_zero");
}
int honei_hip_fill(int vt, void *x, double value, size_t n, void *stream) {
    if (!valid_vtype(vt)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (n && !x) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    return n ? check(vfill(x, value, (int64_t)n, (ValueType)vt, S(stream)), "fill") : HONEI_HIP_SUCCESS;
}
int honei_hip_synchronize(void *stream) {
    return check(stream ? hipStreamSynchronize(S(stream)) : hipDeviceSynchronize(), "synchronize");
}
int honei_hip_stream_create(void **stream) {
    if (!stream) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "stream is null");
    hipStream_t s;
    int st = check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    if (!st) *stream = (void *)s;
    return st;
}
int honei_hip_stream_destroy(void *stream) { return stream ? check(hipStreamDestroy(S(stream)), "hipStreamDestroy") : 0; }

int honei_hip_event_create(void **event) {
    if (!event) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "event is null");
    hipEvent_t e;
    int st = check(hipEventCreate(&e), "hipEventCreate");
    if (!st) *event = (void *)e;
    return st;
}
int honei_hip_event_destroy(void *event) {
    return event ? check(hipEventDestroy((hipEvent_t)event), "hipEventDestroy") : HONEI_HIP_SUCCESS;
}
int honei_hip_event_record(void *event, void *stream) {
    if (!event) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "event is null");
    return check(hipEventRecord((hipEvent_t)event, S(stream)), "hipEventRecord");
}
int honei_hip_event_elapsed_ms(float *ms, void *start, void *stop) {
    if (!ms || !start || !stop) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    int st = check(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize");
    if (st) return st;
    return check(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}

int honei_hip_convert_double_float(float *dst, const double *src, size_t n, void *stream) {
    if (n && (!dst || !src)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    return n ? check(convert_d2f(dst, src, (int64_t)n, S(stream)), "convert_double_float") : 0;
}
int honei_hip_convert_float_double(double *dst, const float *src, size_t n, void *stream) {
    if (n && (!dst || !src)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    return n ? check(convert_f2d(dst, src, (int64_t)n, S(stream)), "convert_float_double") : 0;
}

int honei_hip_upload_indices(void *dst, const uint64_t *src, size_t n, uint64_t bound) {
    if (n == 0) return HONEI_HIP_SUCCESS;
    if (!dst || !src) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    std::vector<uint32_t> tmp(n);
    for (size_t i = 0; i < n; ++i) {
        if (src[i] > 0xffffffffull || (bound != UINT64_MAX && src[i] >= bound))
            return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT,
                        "index " + std::to_string(src[i]) + " at " + std::to_string(i) + " out of range");
        tmp[i] = (uint32_t)src[i];
    }
    return check(hipMemcpy(dst, tmp.data(), n * 4, hipMemcpyHostToDevice), "upload_indices");
}

int honei_hip_upload_values_bf16(void *dst, const float *src, size_t n) {
    if (n == 0) return HONEI_HIP_SUCCESS;
    if (!dst || !src) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    std::vector<uint16_t> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = f32_to_bf16_rne(src[i]);
    return check(hipMemcpy(dst, tmp.data(), n * 2, hipMemcpyHostToDevice), "upload_values_bf16");
}

// ---- ELL ------------------------------------------------------------------
#define PRODUCT_ENTRY(NAME, VT)                                                                          \
    int NAME(void *x, void *y, const void *Aj, const void *Ax, const void *Arl, size_t row_start,          \
             size_t row_end, size_t ncpr, size_t stride, size_t blocksize, size_t threads, void *stream) {   \
        if (row_end > row_start && (!x || !y)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null vector");\
        EllArgs a;                                                                                        \
        a.mode = ELL_PRODUCT;                                                                             \
        a.out = y;                                                                                        \
        a.x = x;                                                                                          \
        a.row_start = (int64_t)row_start;                                                                 \
        a.row_end = (int64_t)row_end;                                                                     \
        return ell_call(VT, a, Aj, Ax, Arl, ncpr, stride, threads, blocksize, stream, #NAME);             \
    }
PRODUCT_ENTRY(honei_hip_product_smell_dv_float, VT_F32)
PRODUCT_ENTRY(honei_hip_product_smell_dv_double, VT_F64)
PRODUCT_ENTRY(honei_hip_product_smell_dv_bf16, VT_BF16)

#define DEFECT_ENTRY(NAME, VT)                                                                            \
    int NAME(void *rhs, void *result, const void *Aj, const void *Ax, const void *Arl, void *b,            \
             size_t row_start, size_t row_end, size_t ncpr, size_t stride, size_t blocksize, size_t threads, \
             void *stream) {                                                                               \
        if (row_end > row_start && (!rhs || !result || !b))                                               \
            return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null vector");                                  \
        EllArgs a;                                                                                         \
        a.mode = ELL_DEFECT;                                                                               \
        a.out = result;                                                                                    \
        a.x = b;                                                                                           \
        a.rhs = rhs;                                                                                       \
        a.row_start = (int64_t)row_start;                                                                  \
        a.row_end = (int64_t)row_end;                                                                      \
        return ell_call(VT, a, Aj, Ax, Arl, ncpr, stride, threads, blocksize, stream, #NAME);              \
    }
DEFECT_ENTRY(honei_hip_defect_smell_dv_float, VT_F32)
DEFECT_ENTRY(honei_hip_defect_smell_dv_double, VT_F64)

// The same two operations on the reference's own index arrays: the glue
// (la/product-cuda.cc:178-194, math/defect-cuda.cc) hands the launcher the
// device copies of SparseMatrixELL::Aj() and Arl(), 64-bit unsigned long
// arrays, which these entries read as they are (low word of each index).
#define PRODUCT_ENTRY64(NAME, VT)                                                                        \
    int NAME(void *x, void *y, const void *Aj, const void *Ax, const void *Arl, size_t row_start,          \
             size_t row_end, size_t ncpr, size_t stride, size_t blocksize, size_t threads, void *stream) {   \
        if (row_end > row_start && (!x || !y)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null vector");\
        EllArgs a;                                                                                        \
        a.mode = ELL_PRODUCT;                                                                             \
        a.out = y;                                                                                        \
        a.x = x;                                                                                          \
        a.row_start = (int64_t)row_start;                                                                 \
        a.row_end = (int64_t)row_end;                                                                     \
        return ell_call(VT, a, Aj, Ax, Arl, ncpr, stride, threads, blocksize, stream, #NAME, 2);          \
    }
PRODUCT_ENTRY64(honei_hip_product_smell_dv_float_idx64, VT_F32)
PRODUCT_ENTRY64(honei_hip_product_smell_dv_double_idx64, VT_F64)

#define DEFECT_ENTRY64(NAME, VT)                                                                          \
    int NAME(void *rhs, void *result, const void *Aj, const void *Ax, const void *Arl, void *b,            \
             size_t row_start, size_t row_end, size_t ncpr, size_t stride, size_t blocksize, size_t threads, \
             void *stream) {                                                                               \
        if (row_end > row_start && (!rhs || !result || !b))                                               \
            return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null vector");                                  \
        EllArgs a;                                                                                         \
        a.mode = ELL_DEFECT;                                                                               \
        a.out = result;                                                                                    \
        a.x = b;                                                                                           \
        a.rhs = rhs;                                                                                       \
        a.row_start = (int64_t)row_start;                                                                  \
        a.row_end = (int64_t)row_end;                                                                      \
        return ell_call(VT, a, Aj, Ax, Arl, ncpr, stride, threads, blocksize, stream, #NAME, 2);           \
    }
DEFECT_ENTRY64(honei_hip_defect_smell_dv_float_idx64, VT_F32)
DEFECT_ENTRY64(honei_hip_defect_smell_dv_double_idx64, VT_F64)

// ---- sliced ELL copy ----------------------------------------------------------
struct SellHandle {
    ValueType vt;
    int64_t rows = 0, cols = 0, slices = 0, padded = 0;
    int32_t *perm = nullptr, *width = nullptr, *Aj = nullptr;
    uint32_t *len = nullptr;
    int64_t *base = nullptr;
    void *Ax = nullptr;
    uint64_t bytes = 0;
    int32_t *ranges = nullptr;   // column range of each sort window
    int64_t window = 0;          // widest window (entries of x), 0 if none fits LDS
    int stage = 1;               // gather x from LDS where the window fits
    ~SellHandle() {
        for (void *p : {(void *)perm, (void *)width, (void *)Aj, (void *)len, (void *)base, Ax, (void *)ranges})
            if (p) (void)hipFree(p);
    }
};

int honei_hip_sell_create(void **sell, int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                          size_t cols, size_t ncpr, size_t stride, size_t threads, void *stream) {
    if (!sell) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "sell is null");
    *sell = nullptr;
    if (vtype != HONEI_HIP_F32 && vtype != HONEI_HIP_F64) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "FP32 or FP64 values");
    if (threads != 1) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "the sliced copy is made from a threads = 1 matrix");
    int st = check_ell(Aj, Ax, Arl, 0, rows, stride, threads, 0);
    if (st) return st;
    if (cols > (size_t)INT32_MAX) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "columns exceed 32-bit index range");
    hipStream_t s = S(stream);
    auto *h = new (std::nothrow) SellHandle();
    if (!h) return fail(HONEI_HIP_ERROR_OUT_OF_MEMORY, "host allocation");
    h->vt = (ValueType)vtype;
    h->rows = (int64_t)rows;
    h->cols = (int64_t)cols;
    const int64_t C = sell_slice_rows();
    h->slices = (h->rows + C - 1) / C;
    const size_t vs = vtype == HONEI_HIP_F64 ? 8 : 4;
    auto dalloc = [&](void **p, size_t b) -> int {
        h->bytes += b ? b : 1;
        return check(hipMalloc(p, b ? b : 1), "sell allocation");
    };
    st = dalloc((void **)&h->perm, rows * 4);
    if (!st) st = dalloc((void **)&h->len, rows * 4);
    if (!st) st = dalloc((void **)&h->width, (size_t)h->slices * 4);
    if (!st) st = dalloc((void **)&h->base, (size_t)h->slices * 8);
    if (!st) st = check(sell_sort((const uint32_t *)Arl, h->rows, h->perm, h->len, s), "sell sort");
    if (!st) st = check(sell_widths(h->len, h->rows, h->slices, h->width, s), "sell widths");
    std::vector<int32_t> w((size_t)h->slices);
    std::vector<int64_t> b((size_t)h->slices);
    if (!st && h->slices)
        st = check(hipMemcpyAsync(w.data(), h->width, w.size() * 4, hipMemcpyDeviceToHost, s), "sell widths readback");
    if (!st) st = check(hipStreamSynchronize(s), "sell sync");
    int64_t acc = 0;
    for (int64_t k = 0; !st && k < h->slices; ++k) {
        if (w[(size_t)k] < 0 || (size_t)w[(size_t)k] > ncpr) st = fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "row length exceeds num_cols_per_row");
        b[(size_t)k] = acc;
        acc += C * (int64_t)w[(size_t)k];
    }
    h->padded = acc;
    if (!st && h->slices)
        st = check(hipMemcpyAsync(h->base, b.data(), b.size() * 8, hipMemcpyHostToDevice, s), "sell base upload");
    if (!st) st = dalloc((void **)&h->Aj, (size_t)acc * 4);
    if (!st) st = dalloc(&h->Ax, (size_t)acc * vs);
    if (!st)
        st = check(sell_fill(h->vt, (const int32_t *)Aj, Ax, (int64_t)stride, h->rows, h->perm, h->len, h->width,
                             h->base, h->Aj, h->Ax, s),
                   "sell fill");
    // each sort window's column range, for staging x in LDS
    const int64_t windows = (h->rows + sell_window() - 1) / sell_window();
    std::vector<int32_t> rg((size_t)(2 * windows));
    if (!st && windows) st = dalloc((void **)&h->ranges, (size_t)windows * 8);
    if (!st && windows)
        st = check(sell_ranges(h->Aj, h->len, h->width, h->base, h->rows, h->ranges, s), "sell ranges");
    if (!st && windows)
        st = check(hipMemcpyAsync(rg.data(), h->ranges, rg.size() * 4, hipMemcpyDeviceToHost, s), "sell ranges readback");
    if (!st) st = check(hipStreamSynchronize(s), "sell sync");
    if (!st) {
        const int64_t cap = (int64_t)(kMaxStageBytes / vs);
        for (int64_t k = 0; k < windows; ++k) {
            const int64_t wn = (int64_t)rg[(size_t)(2 * k + 1)] - rg[(size_t)(2 * k)] + 1;
            if (wn > h->window) h->window = wn;
        }
        if (h->window > cap) h->window = cap;   // wider windows gather from global x
    }
    if (st) {
        delete h;
        return st;
    }
    *sell = h;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_sell_info(void *sell, size_t *slices, size_t *padded_slots, size_t *device_bytes) {
    if (!sell) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "sell is null");
    const SellHandle *h = (const SellHandle *)sell;
    if (slices) *slices = (size_t)h->slices;
    if (padded_slots) *padded_slots = (size_t)h->padded;
    if (device_bytes) *device_bytes = h->bytes;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_product_sell(void *sell, void *x, void *y, void *stream) {
    if (!sell) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "sell is null");
    const SellHandle *h = (const SellHandle *)sell;
    if (h->rows > 0 && (!x || !y)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null vector");
    if (h->stage && h->window > 0)
        return check(sell_spmv_staged(h->vt, h->Aj, h->Ax, h->perm, h->len, h->width, h->base, h->rows, x, y,
                                      h->ranges, h->window, S(stream)),
                     "product_sell (staged)");
    return check(sell_spmv(h->vt, h->Aj, h->Ax, h->perm, h->len, h->width, h->base, h->rows, x, y, S(stream)),
                 "product_sell");
}

int honei_hip_sell_stage_x(void *sell, int on) {
    if (!sell) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "sell is null");
    if (on != 0 && on != 1) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "on must be 0 or 1");
    ((SellHandle *)sell)->stage = on;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_sell_destroy(void *sell) {
    delete (SellHandle *)sell;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_stream_probe(int kind, void *dst, const void *src, size_t bytes, size_t grid, void *stream) {
    if (kind < 0 || kind > 3)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "kind: 0 read, 1 copy (grid-stride), 2 read, 3 copy (chunked)");
    if (!dst || !src) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bytes and pointers must be multiples of 16");
    if (grid < 1 || grid > 65536) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "grid outside 1..65536");
    return check(stream_probe(kind, dst, src, (int64_t)(bytes / 16), (int)grid, S(stream)), "stream_probe");
}

int honei_hip_narrow_indices(void *dst, const void *src, size_t n, uint64_t bound, void *stream) {
    if (n == 0) return HONEI_HIP_SUCCESS;
    if (!dst || !src) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    if (n > (size_t)INT64_MAX / 8) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "n too large");
    Workspace *w = workspace_for(S(stream));
    if (!w) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    const uint64_t lim = bound < (uint64_t(1) << 32) ? bound : (uint64_t(1) << 32);
    unsigned int *bad = (unsigned int *)(w->scalars + Workspace::kScalars - 1);
    int st = check(fill_bytes_zero(bad, sizeof(unsigned int), S(stream)), "narrow_indices clear");
    if (!st)
        st = check(narrow_indices((uint32_t *)dst, (const uint64_t *)src, (int64_t)n, lim, bad, S(stream)),
                   "narrow_indices");
    unsigned int hbad = 0;
    if (!st) st = check(hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, S(stream)), "narrow readback");
    if (!st) st = check(hipStreamSynchronize(S(stream)), "narrow sync");
    if (st) return st;
    if (hbad)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT,
                    std::to_string(hbad) + " indices at or above " + std::to_string(lim));
    return HONEI_HIP_SUCCESS;
}

// ---- ELL product with x staged in LDS -------------------------------------------
int honei_hip_ell_block_ranges(const void *Aj, const void *Arl, size_t rows, size_t ncpr, size_t stride,
                               size_t threads, size_t rows_per_block, int32_t *ranges, int64_t *max_width,
                               void *stream) {
    int st = check_ell(Aj, Aj, Arl, 0, rows, stride, threads, 0);
    if (st) return st;
    if (!ranges || rows_per_block < 1) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "ranges / rows_per_block");
    EllView A = make_view(Aj, Aj, Arl, rows, ncpr, stride, threads, VT_F32);
    st = check(ell_block_ranges(A, (int64_t)rows_per_block, ranges, S(stream)), "ell_block_ranges");
    if (st || !max_width) return st;
    const size_t nb = (rows + rows_per_block - 1) / rows_per_block;
    std::vector<int32_t> h(2 * nb);
    st = check(hipMemcpyAsync(h.data(), ranges, h.size() * 4, hipMemcpyDeviceToHost, S(stream)), "ranges readback");
    if (!st) st = check(hipStreamSynchronize(S(stream)), "ranges sync");
    if (st) return st;
    int64_t w = 0;
    for (size_t b = 0; b < nb; ++b)
        if (h[2 * b + 1] >= h[2 * b]) w = std::max<int64_t>(w, (int64_t)h[2 * b + 1] - h[2 * b] + 1);
    *max_width = w;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_product_smell_dv_staged(int vtype, void *x, void *y, const void *Aj, const void *Ax, const void *Arl,
                                      size_t rows, size_t ncpr, size_t stride, size_t threads, size_t blocksize,
                                      size_t chunks, const int32_t *ranges, size_t nranges, int64_t window,
                                      void *stream) {
    int st = check_ell(Aj, Ax, Arl, 0, rows, stride, threads, blocksize);
    if (st) return st;
    if (rows > 0 && (!x || !y || !ranges)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null vector or ranges");
    if (chunks < 1 || chunks > 64) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "chunks must be 1 .. 64");
    const size_t bs = blocksize ? blocksize : 256, rpb = bs / threads * chunks;
    if (nranges < (rows + rpb - 1) / rpb)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "fewer ranges than blocks: make them with rows_per_block = "
                                                      "chunks * blocksize / threads");
    if (vtype < VT_F32 || vtype > VT_BF16) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "vtype");
    const size_t vs = vtype == VT_F64 ? 8 : 4;
    if (window < 1 || (size_t)window * vs > kMaxStageBytes)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "window must be 1 .. 96 KiB of vector elements");
    EllView A = make_view(Aj, Ax, Arl, rows, ncpr, stride, threads, (ValueType)vtype);
    return check(ell_product_staged(A, y, x, ranges, (int)bs, (int)chunks, window, S(stream)), "product_smell_dv_staged");
}

// ---- banded Q1 ------------------------------------------------------------------
static int banded_call(ValueType vt, void *rhs, void *ll, void *ld, void *lu, void *dl, void *dd, void *du, void *ul,
                       void *ud, void *uu, void *x, void *y, size_t size, size_t blocksize, size_t m, void *stream,
                       const char *what) {
    if (size == 0) return HONEI_HIP_SUCCESS;
    if (!x || !y) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null vector");
    if (m == 0 || m * m != size) return fail(HONEI_HIP_ERROR_SIZE_MISMATCH, "size must be m*m");
    if (blocksize && (blocksize % 64 || blocksize > 256))
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "blocksize: 0 or a multiple of 64 up to 256");
    const void *bands[9] = {ll, ld, lu, dl, dd, du, ul, ud, uu};
    for (const void *b : bands)
        if (!b) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null band");
    return check(q1_banded(vt, bands, x, rhs, y, (int64_t)size, (int64_t)m, (int)blocksize, S(stream)), what);
}

#define BANDED_PRODUCT_ENTRY(NAME, VT)                                                                    \
    int NAME(void *ll, void *ld, void *lu, void *dl, void *dd, void *du, void *ul, void *ud, void *uu,     \
             void *x, void *y, size_t size, size_t blocksize, size_t m, void *stream) {                    \
        return banded_call(VT, nullptr, ll, ld, lu, dl, dd, du, ul, ud, uu, x, y, size, blocksize, m, stream, \
                           #NAME);                                                                         \
    }
#define BANDED_DEFECT_ENTRY(NAME, VT)                                                                     \
    int NAME(void *rhs, void *ll, void *ld, void *lu, void *dl, void *dd, void *du, void *ul, void *ud,    \
             void *uu, void *x, void *y, size_t size, size_t blocksize, size_t m, void *stream) {          \
        if (size && !rhs) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null rhs");                       \
        return banded_call(VT, rhs, ll, ld, lu, dl, dd, du, ul, ud, uu, x, y, size, blocksize, m, stream,  \
                           #NAME);                                                                         \
    }
BANDED_PRODUCT_ENTRY(honei_hip_product_bmdv_q1_float, VT_F32)
BANDED_PRODUCT_ENTRY(honei_hip_product_bmdv_q1_double, VT_F64)
BANDED_DEFECT_ENTRY(honei_hip_defect_q1_float, VT_F32)
BANDED_DEFECT_ENTRY(honei_hip_defect_q1_double, VT_F64)

int honei_hip_defect_norm_smell_dv(int vtype, void *rhs, void *result, const void *Aj, const void *Ax,
                                   const void *Arl, void *x, size_t rows, size_t ncpr, size_t stride, size_t threads,
                                   double *norm2_dev, void *stream) {
    if (vtype != HONEI_HIP_F32 && vtype != HONEI_HIP_F64 && vtype != HONEI_HIP_BF16)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (rows && (!rhs || !result || !x || !norm2_dev)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    EllArgs a;
    a.mode = ELL_DEFECT;
    a.out = result;
    a.x = x;
    a.rhs = rhs;
    a.norm2 = norm2_dev;
    a.row_start = 0;
    a.row_end = (int64_t)rows;
    return ell_call((ValueType)vtype, a, Aj, Ax, Arl, ncpr, stride, threads, 0, stream, "defect_norm");
}

int honei_hip_jacobi_smell_dv(int vtype, void *x_new, const void *x, const void *rhs, const void *dinv, const void *Aj,
                              const void *Ax, const void *Arl, size_t rows, size_t ncpr, size_t stride, size_t threads,
                              void *acc, void *stream) {
    if (vtype != HONEI_HIP_F32 && vtype != HONEI_HIP_F64 && vtype != HONEI_HIP_BF16)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (rows && (!x_new || !rhs || !dinv)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    if (x && x == x_new) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "x_new must not alias x");
    EllArgs a;
    a.mode = x ? ELL_JACOBI : ELL_JACOBI0;
    a.out = x_new;
    a.x = x;
    a.rhs = rhs;
    a.dinv = dinv;
    a.acc = acc;
    a.row_start = 0;
    a.row_end = (int64_t)rows;
    return ell_call((ValueType)vtype, a, Aj, Ax, Arl, ncpr, stride, threads, 0, stream, "jacobi");
}

// ---- BLAS-1 ---------------------------------------------------------------
static int host_reduce(double *out, int vt, const void *x, const void *y, size_t n, void *stream, bool is_dot) {
    if (n && (!x || (is_dot && !y))) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null vector");
    RedCtx c{x, y, (int64_t)n, (ValueType)vt, S(stream)};
    return reduce_to_host(out, stream, is_dot ? launch_dot : launch_norm, &c);
}
int honei_hip_dot_product_two_float(float *result, const void *x, const void *y, size_t n, void *stream) {
    double v;
    int st = host_reduce(&v, VT_F32, x, y, n, stream, true);
    if (!st && result) *result = (float)v;
    return st;
}
int honei_hip_dot_product_two_double(double *result, const void *x, const void *y, size_t n, void *stream) {
    double v;
    int st = host_reduce(&v, VT_F64, x, y, n, stream, true);
    if (!st && result) *result = v;
    return st;
}
int honei_hip_norm_l2_one_float(float *result, const void *x, size_t n, void *stream) {
    double v;
    int st = host_reduce(&v, VT_F32, x, x, n, stream, false);
    if (!st && result) *result = (float)v;
    return st;
}
int honei_hip_norm_l2_one_double(double *result, const void *x, size_t n, void *stream) {
    double v;
    int st = host_reduce(&v, VT_F64, x, x, n, stream, false);
    if (!st && result) *result = v;
    return st;
}
int honei_hip_dot_product_dev(int vt, double *out, const void *x, const void *y, size_t n, void *stream) {
    if (!valid_vtype(vt)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!out || (n && (!x || !y))) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    Workspace *w = workspace_for(S(stream));
    if (!w) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    return check(dot(x, y, (int64_t)n, (ValueType)vt, out, w->slots, S(stream)), "dot");
}
int honei_hip_norm_l2_dev(int vt, double *out, const void *x, size_t n, void *stream) {
    if (!valid_vtype(vt)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!out || (n && !x)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    Workspace *w = workspace_for(S(stream));
    if (!w) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    return check(norm2(x, (int64_t)n, (ValueType)vt, out, w->slots, S(stream)), "norm");
}
int honei_hip_axpy_norm_dev(int vt, void *x, const void *y, double alpha, double *out, size_t n, void *stream) {
    if (!valid_vtype(vt)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!out || (n && (!x || !y))) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    Workspace *w = workspace_for(S(stream));
    if (!w) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    return check(axpy_norm2(x, y, alpha, nullptr, (int64_t)n, (ValueType)vt, out, w->slots, S(stream)), "axpy_norm");
}

#define VEC_GUARD(...)                                                                  \
    if (!valid_vtype(vt)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");   \
    if (n == 0) return HONEI_HIP_SUCCESS;                                               \
    {                                                                                   \
        const void *ps[] = {__VA_ARGS__};                                               \
        for (const void *p : ps)                                                        \
            if (!p) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null vector");       \
    }

int honei_hip_scaled_sum_two(int vt, void *x, const void *y, double s, size_t n, void *stream) {
    VEC_GUARD(x, y);
    return check(axpy(x, y, s, nullptr, (int64_t)n, (ValueType)vt, S(stream)), "scaled_sum_two");
}
int honei_hip_scaled_sum_three_s(int vt, void *r, const void *x, const void *y, double s, size_t n, void *stream) {
    VEC_GUARD(r, x, y);
    return check(scaled_sum3(r, x, y, s, (int64_t)n, (ValueType)vt, S(stream)), "scaled_sum_three_s");
}
int honei_hip_scaled_sum_three(int vt, void *x, const void *y, const void *z, size_t n, void *stream) {
    VEC_GUARD(x, y, z);
    return check(scaled_sum_v(x, y, z, (int64_t)n, (ValueType)vt, S(stream)), "scaled_sum_three");
}
int honei_hip_sum_two(int vt, void *a, const void *b, size_t n, void *stream) {
    VEC_GUARD(a, b);
    return check(vsum(a, b, (int64_t)n, (ValueType)vt, S(stream)), "sum_two");
}
int honei_hip_difference_two(int vt, void *a, const void *b, size_t n, void *stream) {
    VEC_GUARD(a, b);
    return check(vdiff2(a, b, (int64_t)n, (ValueType)vt, S(stream)), "difference_two");
}
int honei_hip_difference_three(int vt, void *r, const void *a, const void *b, size_t n, void *stream) {
    VEC_GUARD(r, a, b);
    return check(vdiff3(r, a, b, (int64_t)n, (ValueType)vt, S(stream)), "difference_three");
}
int honei_hip_scale_one(int vt, void *x, double a, size_t n, void *stream) {
    VEC_GUARD(x);
    return check(vscale(x, a, (int64_t)n, (ValueType)vt, S(stream)), "scale_one");
}
int honei_hip_element_product_three(int vt, void *r, const void *a, const void *b, size_t n, void *stream) {
    VEC_GUARD(r, a, b);
    return check(vprod(r, a, b, (int64_t)n, (ValueType)vt, S(stream)), "element_product_three");
}
int honei_hip_element_inverse_one(int vt, void *x, size_t n, void *stream) {
    VEC_GUARD(x);
    return check(vinverse(x, (int64_t)n, (ValueType)vt, S(stream)), "element_inverse_one");
}

// ---- transfers --------------------------------------------------------------
int honei_hip_restriction(int vt, void *coarse, size_t nc, const void *fine, size_t nf, const uint64_t *mask,
                          void *stream) {
    if (!valid_vtype(vt)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!coarse || !fine || !mask) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    hipError_t e = restriction(coarse, (int64_t)nc, fine, (int64_t)nf, mask, (ValueType)vt, S(stream));
    if (e == hipErrorInvalidValue)
        return fail(HONEI_HIP_ERROR_SIZE_MISMATCH, "sizes are not (2^L+1)^2 fine / (2^(L-1)+1)^2 coarse");
    return check(e, "restriction");
}
int honei_hip_prolongation(int vt, void *fine, size_t nf, const void *coarse, size_t nc, const uint64_t *mask,
                           void *stream) {
    if (!valid_vtype(vt)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!coarse || !fine || !mask) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    hipError_t e = prolongation(fine, (int64_t)nf, coarse, (int64_t)nc, mask, (ValueType)vt, S(stream));
    if (e == hipErrorInvalidValue)
        return fail(HONEI_HIP_ERROR_SIZE_MISMATCH, "sizes are not (2^L+1)^2 fine / (2^(L-1)+1)^2 coarse");
    return check(e, "prolongation");
}

}  // extern "C"
