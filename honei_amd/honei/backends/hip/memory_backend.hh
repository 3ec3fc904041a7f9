// honei/backends/hip/memory_backend.hh -- MemoryBackend<tags::GPU::HIP>, the
// device memory space of the MI355X backend behind HONEI's MemoryArbiter
// interface (util/memory_backend_base.hh:32-58), in the role of
// MemoryBackend<tags::GPU::CUDA> (backends/cuda/memory_backend.cc:209-470).
//
// A block is keyed by its memid (the host array's id) and the host address
// it mirrors.  Bytes are copied as they are: the device copy of an
// unsigned long array (SparseMatrixELL::Aj(), Arl()) stays 64-bit, and the
// operator glue passes it to the *_idx64 entries of include/honei_hip.h,
// which read that layout directly.  One device; every transfer is
// synchronous on the default stream, as the arbiter expects a block to be
// usable when the call returns.  The arbiter serialises calls under its own
// lock; the backend keeps its map under a mutex of its own as well.
//
// To register it with the reference's arbiter a maintainer adds, next to
// the CUDA one (backends/cuda/memory_backend.cc:470-480):
//   static MemoryBackendRegistrator hip_backend_registrator(
//       tags::GPU::HIP::memory_value, MemoryBackend<tags::GPU::HIP>::instance());
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "honei/util/exception.hh"
#include "honei/util/memory_backend_base.hh"
#include "honei/util/tags.hh"
#include "honei_hip.h"

namespace honei {

template <typename Tag_> class MemoryBackend;

template <>
class MemoryBackend<tags::GPU::HIP> : public MemoryBackendBase {
    struct Chunk {
        void *address;
        void *device;
        unsigned long bytes;
    };
    std::multimap<void *, Chunk> _id_map;       // memid -> its device chunks
    std::map<void *, void *> _address_map;      // host address -> device copy
    std::mutex _mutex;

    Chunk *find(void *memid, void *address) {
        auto range = _id_map.equal_range(memid);
        for (auto i = range.first; i != range.second; ++i)
            if (i->second.address == address) return &i->second;
        return nullptr;
    }
    void *device_of(void *address, const char *what) {
        auto i = _address_map.find(address);
        if (i == _address_map.end()) throw InternalError(std::string("MemoryBackend<tags::GPU::HIP>::") + what +
                                                         " address not found!");
        return i->second;
    }
    void *alloc_locked(void *memid, void *address, unsigned long bytes) {
        if (Chunk *c = find(memid, address)) return c->device;
        void *device = nullptr;
        hip_check(honei_hip_malloc(&device, bytes ? bytes : 1), "MemoryBackend<tags::GPU::HIP>::alloc");
        _id_map.insert(std::make_pair(memid, Chunk{address, device, bytes}));
        _address_map[address] = device;
        return device;
    }

  public:
    static MemoryBackend *instance() {
        static MemoryBackend backend;
        return &backend;
    }
    ~MemoryBackend() override {
        for (auto &kv : _id_map) honei_hip_free(kv.second.device);
    }

    // Block count and bytes held (for tests and diagnostics).
    unsigned long blocks() {
        std::lock_guard<std::mutex> l(_mutex);
        return (unsigned long)_id_map.size();
    }

    void *upload(void *memid, void *address, unsigned long bytes) override {
        std::lock_guard<std::mutex> l(_mutex);
        // the arbiter calls upload only while this memory still holds a
        // valid copy or has none (it frees stale copies on writes elsewhere)
        if (Chunk *c = find(memid, address)) return c->device;
        void *device = alloc_locked(memid, address, bytes);
        if (bytes) hip_check(honei_hip_upload(device, address, bytes, nullptr), "MemoryBackend<tags::GPU::HIP>::upload");
        return device;
    }

    void download(void *memid, void *, unsigned long) override {
        std::lock_guard<std::mutex> l(_mutex);
        auto range = _id_map.equal_range(memid);
        if (range.first == range.second)
            throw InternalError("MemoryBackend<tags::GPU::HIP>::download address not found!");
        for (auto i = range.first; i != range.second; ++i)
            if (i->second.bytes)
                hip_check(honei_hip_download(i->second.address, i->second.device, i->second.bytes, nullptr),
                          "MemoryBackend<tags::GPU::HIP>::download");
    }

    void *alloc(void *memid, void *address, unsigned long bytes) override {
        std::lock_guard<std::mutex> l(_mutex);
        return alloc_locked(memid, address, bytes);
    }

    void free(void *memid) override {
        std::lock_guard<std::mutex> l(_mutex);
        auto range = _id_map.equal_range(memid);
        for (auto i = range.first; i != range.second; ++i) {
            hip_check(honei_hip_free(i->second.device), "MemoryBackend<tags::GPU::HIP>::free");
            _address_map.erase(i->second.address);
        }
        _id_map.erase(range.first, range.second);
    }

    void copy(void *, void *src_address, void *dest_id, void *dest_address, unsigned long bytes) override {
        std::lock_guard<std::mutex> l(_mutex);
        void *src = device_of(src_address, "copy");
        void *dst = device_of(dest_address, "copy");
        Chunk *d = find(dest_id, dest_address);
        if (d && d->bytes < bytes) throw InternalError("MemoryBackend<tags::GPU::HIP>::copy destination too small");
        if (bytes) hip_check(honei_hip_copy(dst, src, bytes, nullptr), "MemoryBackend<tags::GPU::HIP>::copy");
        hip_check(honei_hip_synchronize(nullptr), "MemoryBackend<tags::GPU::HIP>::copy");
    }

    void convert_float_double(void *, void *src_address, void *, void *dest_address, unsigned long bytes) override {
        std::lock_guard<std::mutex> l(_mutex);
        float *src = (float *)device_of(src_address, "convert_float_double");
        double *dst = (double *)device_of(dest_address, "convert_float_double");
        hip_check(honei_hip_convert_float_double(dst, src, bytes / sizeof(float), nullptr),
                  "MemoryBackend<tags::GPU::HIP>::convert_float_double");
        hip_check(honei_hip_synchronize(nullptr), "MemoryBackend<tags::GPU::HIP>::convert_float_double");
    }

    void convert_double_float(void *, void *src_address, void *, void *dest_address, unsigned long bytes) override {
        std::lock_guard<std::mutex> l(_mutex);
        double *src = (double *)device_of(src_address, "convert_double_float");
        float *dst = (float *)device_of(dest_address, "convert_double_float");
        hip_check(honei_hip_convert_double_float(dst, src, bytes / sizeof(double), nullptr),
                  "MemoryBackend<tags::GPU::HIP>::convert_double_float");
        hip_check(honei_hip_synchronize(nullptr), "MemoryBackend<tags::GPU::HIP>::convert_double_float");
    }

    // Any value (the CUDA backend accepts only zero).
    void fill(void *, void *address, unsigned long bytes, float proto) override {
        std::lock_guard<std::mutex> l(_mutex);
        void *dev = device_of(address, "fill");
        hip_check(honei_hip_fill(HONEI_HIP_F32, dev, (double)proto, bytes / sizeof(float), nullptr),
                  "MemoryBackend<tags::GPU::HIP>::fill");
        hip_check(honei_hip_synchronize(nullptr), "MemoryBackend<tags::GPU::HIP>::fill");
    }
    void fill(void *, void *address, unsigned long bytes, double proto) override {
        std::lock_guard<std::mutex> l(_mutex);
        void *dev = device_of(address, "fill");
        hip_check(honei_hip_fill(HONEI_HIP_F64, dev, proto, bytes / sizeof(double), nullptr),
                  "MemoryBackend<tags::GPU::HIP>::fill");
        hip_check(honei_hip_synchronize(nullptr), "MemoryBackend<tags::GPU::HIP>::fill");
    }

    bool knows(void *memid, void *address) override {
        std::lock_guard<std::mutex> l(_mutex);
        return find(memid, address) != nullptr;
    }
};

}  // namespace honei
