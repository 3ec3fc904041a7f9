// noalloc_capi.cpp -- CGSolver on a device ELL matrix with no allocation at
// call time (honei_hip_cg_solve_noalloc, include/honei_hip.h).
//
// The solve's scratch (reduction partials, the scalar record and the vectors
// p, r, v) lives in `scratch`, a file-scope device array that every device
// the library runs on holds from module load until the process ends.  Host
// code never passes the array's name to a kernel: on the host it is only a
// stand-in for the symbol.  The device address is looked up once per device
// with hipGetSymbolAddress, with that device current, and kept in
// g_scratch_ptr[device]; kernels, fills and copies use that address.
//
// Every call on one device shares that device's array and leaves its
// contents behind; the call clears what must start at zero.  Calls that
// overlap on one device (two host threads, or two streams) race on it.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "honei_hip.h"
#include "runtime.hh"
#include "solvers.hh"

using namespace honei::hipx;

namespace {

// Partial slots as the workspace of honei_hip_cg_solve (same capacity, so
// the launches and their summation order are the same), then 4 KiB for the
// counter and the scalar record, then the three vectors.
constexpr size_t kPartialBytes = sizeof(double) * (size_t)Workspace::kSlotCapacity * Workspace::kSlotLanes;
constexpr size_t kCounterOff = kPartialBytes;
constexpr size_t kScalarOff = kPartialBytes + 256;
constexpr size_t kVectorOff = kPartialBytes + 4096;
constexpr size_t kVectorBytes = (size_t)32 << 20;   // p, r and v together
constexpr size_t kScratchBytes = kVectorOff + kVectorBytes;
constexpr int kMaxDevices = 64;

static_assert(sizeof(CGScalars) <= kVectorOff - kScalarOff, "scalar record exceeds its scratch region");

std::mutex g_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace honei {
namespace hipx {

// Device address of `scratch` on the current device, looked up on the first
// call for that device.  Returns a HONEI status code.
int noalloc_setup(void **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(HONEI_HIP_ERROR_RUNTIME, "hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return fail(HONEI_HIP_ERROR_RUNTIME, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(HONEI_HIP_ERROR_RUNTIME, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    if (out) *out = g_scratch_ptr[dev];
    return HONEI_HIP_SUCCESS;
}

size_t noalloc_max_vector_bytes() { return kVectorBytes / 3; }

}  // namespace hipx
}  // namespace honei

extern "C" {

int honei_hip_cg_solve_noalloc(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                               size_t num_cols_per_row, size_t stride, size_t threads, const void *b, void *x,
                               size_t max_iters, double eps_relative, size_t *used_iters, double *final_defect,
                               void *stream) {
    int st = cg_check_args(vtype, Aj, Ax, Arl, rows, stride, threads, b, x);
    if (st) return st;
    const size_t vbytes = cg_vector_bytes((ValueType)vtype, rows);
    if (vbytes > noalloc_max_vector_bytes())
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT,
                    "honei_hip_cg_solve_noalloc: " + std::to_string(rows) + " rows need " + std::to_string(3 * vbytes) +
                        " bytes of vectors, beyond the fixed scratch of " + std::to_string(kVectorBytes) + " bytes");
    void *base_ptr = nullptr;
    if ((st = noalloc_setup(&base_ptr))) return st;
    char *base = (char *)base_ptr;
    hipStream_t s = (hipStream_t)stream;
    // the counter stays zero by contract; the scalar record is written by the
    // solve's first kernels, cleared anyway so no earlier call shows through
    if (fill_bytes_zero(base + kCounterOff, kVectorOff - kCounterOff, s) != hipSuccess)
        return fail(HONEI_HIP_ERROR_RUNTIME, "honei_hip_cg_solve_noalloc: scratch clear failed");
    ReduceSlots slots;
    slots.partials = (double *)base;
    slots.counter = (unsigned int *)(base + kCounterOff);
    slots.capacity = Workspace::kSlotCapacity;
    const EllView A = cg_view(vtype, Aj, Ax, Arl, rows, num_cols_per_row, stride, threads);
    char *v0 = base + kVectorOff;
    return cg_solve_on(A, b, x, max_iters, eps_relative, used_iters, final_defect, s, slots, v0, v0 + vbytes,
                       v0 + 2 * vbytes, (CGScalars *)(base + kScalarOff));
}

}  // extern "C"
