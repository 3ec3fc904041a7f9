// runtime.hh -- per-stream workspaces and error reporting of the backend.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "backend.hh"

namespace honei {
namespace hipx {

// Persistent scratch owned by the library, one per (device, stream): the
// cross-block reduction slots plus a small array of device scalars.  Created
// on first use with hipMalloc, kept until honei_hip_shutdown(); no per-call
// allocation on any path.
struct Workspace {
    ReduceSlots slots;
    double *scalars = nullptr;   // kScalars doubles
    static constexpr int kScalars = 256;
    static constexpr int kSlotCapacity = 1 << 17;   // blocks
    static constexpr int kSlotLanes = 4;
    void *scratch = nullptr;     // grows on demand, never shrinks
    size_t scratch_bytes = 0;
    // Returns a buffer of at least `bytes` (reallocating only when larger
    // than any earlier request), or nullptr on allocation failure.
    void *ensure_scratch(size_t bytes);
};

// Returns nullptr (and records the error) when allocation fails.
Workspace *workspace_for(hipStream_t s);
void release_workspaces();

void set_error(const std::string &msg);
const char *last_error();

}  // namespace hipx
}  // namespace honei
