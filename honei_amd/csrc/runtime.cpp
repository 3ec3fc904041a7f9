// runtime.cpp -- workspace registry and thread-local error text.

#include "runtime.hh"

#include <map>
#include <mutex>
#include <utility>

namespace honei {
namespace hipx {

namespace {
std::mutex g_mutex;
std::map<std::pair<int, hipStream_t>, Workspace *> g_ws;
thread_local std::string g_error;
}  // namespace

void set_error(const std::string &msg) { g_error = msg; }
const char *last_error() { return g_error.c_str(); }

Workspace *workspace_for(hipStream_t s) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("hipGetDevice failed");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    auto key = std::make_pair(dev, s);
    auto it = g_ws.find(key);
    if (it != g_ws.end()) return it->second;
    Workspace *w = new Workspace;
    w->slots.capacity = Workspace::kSlotCapacity;
    hipError_t e1 = hipMalloc((void **)&w->slots.partials,
                              sizeof(double) * (size_t)Workspace::kSlotCapacity * Workspace::kSlotLanes);
    hipError_t e2 = hipMalloc((void **)&w->slots.counter, 64);
    hipError_t e3 = hipMalloc((void **)&w->scalars, sizeof(double) * Workspace::kScalars);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        set_error("workspace allocation failed");
        if (e1 == hipSuccess) (void)hipFree(w->slots.partials);
        if (e2 == hipSuccess) (void)hipFree(w->slots.counter);
        if (e3 == hipSuccess) (void)hipFree(w->scalars);
        delete w;
        return nullptr;
    }
    // The counter must start at zero; every reduction leaves it at zero.
    if (fill_bytes_zero(w->slots.counter, 64, s) != hipSuccess ||
        fill_bytes_zero(w->scalars, sizeof(double) * Workspace::kScalars, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_error("workspace initialisation failed");
        (void)hipFree(w->slots.partials);
        (void)hipFree(w->slots.counter);
        (void)hipFree(w->scalars);
        delete w;
        return nullptr;
    }
    g_ws[key] = w;
    return w;
}

void *Workspace::ensure_scratch(size_t bytes) {
    if (bytes <= scratch_bytes && scratch) return scratch;
    if (scratch) {
        (void)hipDeviceSynchronize();
        (void)hipFree(scratch);
        scratch = nullptr;
        scratch_bytes = 0;
    }
    if (hipMalloc(&scratch, bytes) != hipSuccess) {
        set_error("workspace scratch allocation failed");
        scratch = nullptr;
        return nullptr;
    }
    scratch_bytes = bytes;
    return scratch;
}

void release_workspaces() {
    std::lock_guard<std::mutex> lock(g_mutex);
    for (auto &kv : g_ws) {
        (void)hipFree(kv.second->slots.partials);
        (void)hipFree(kv.second->slots.counter);
        (void)hipFree(kv.second->scalars);
        if (kv.second->scratch) (void)hipFree(kv.second->scratch);
        delete kv.second;
    }
    g_ws.clear();
}

}  // namespace hipx
}  // namespace honei
