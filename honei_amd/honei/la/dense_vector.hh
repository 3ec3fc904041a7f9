// honei/la/dense_vector.hh -- DenseVector<DT> with host/device residency.
//
// Same handle semantics as the reference (la/dense_vector-impl.hh:251-272):
// copies share storage, copy() is deep.  lock(mode, memory) returns a
// pointer valid in the requested memory and moves the data there on demand
// -- the role of MemoryArbiter + MemoryBackend<GPU::CUDA>
// (util/memory_arbiter-impl.hh:471-612, backends/cuda/memory_backend.cc:
// 209-470) collapsed into the vector: one device, stream-ordered, so no
// arbiter thread or condition variable is needed.
#pragma once

#include <cstring>
#include <memory>
#include <vector>

#include "honei/util/exception.hh"
#include "honei/util/tags.hh"
#include "honei_hip.h"

namespace honei {

template <typename DT>
class DenseVector {
    struct Imp {
        std::vector<DT> host;
        void *dev = nullptr;
        bool host_valid = true;
        bool dev_valid = false;
        explicit Imp(unsigned long n) : host(n) {}
        ~Imp() {
            if (dev) honei_hip_free(dev);
        }
    };
    std::shared_ptr<Imp> _imp;

    void ensure_dev() const {
        if (!_imp->dev) hip_check(honei_hip_malloc(&_imp->dev, sizeof(DT) * (_imp->host.size() ? _imp->host.size() : 1)),
                                  "DenseVector device allocation");
    }

  public:
    typedef DT DataType;

    explicit DenseVector(unsigned long size) : _imp(std::make_shared<Imp>(size)) {}
    DenseVector(unsigned long size, DT value) : _imp(std::make_shared<Imp>(size)) {
        for (auto &v : _imp->host) v = value;
    }
    DenseVector(const DenseVector &other) = default;
    DenseVector &operator=(const DenseVector &other) = default;

    unsigned long size() const { return _imp->host.size(); }

    // Pointer in `memory`; data moved there unless the lock only writes.
    void *lock(LockMode mode, tags::TagValue memory = tags::tv_cpu) const {
        if (memory == tags::tv_gpu_hip) {
            ensure_dev();
            if (mode != lm_write_only && !_imp->dev_valid && size()) {
                hip_check(honei_hip_upload(_imp->dev, _imp->host.data(), sizeof(DT) * size(), nullptr),
                          "DenseVector upload");
            }
            _imp->dev_valid = true;
            if (mode != lm_read_only) _imp->host_valid = false;
            return _imp->dev;
        }
        if (mode != lm_write_only && !_imp->host_valid && size()) {
            hip_check(honei_hip_download(_imp->host.data(), _imp->dev, sizeof(DT) * size(), nullptr),
                      "DenseVector download");
        }
        _imp->host_valid = true;
        if (mode != lm_read_only) _imp->dev_valid = false;
        return _imp->host.data();
    }
    void unlock(LockMode) const {}

    // Host access (synchronises from the device, invalidates the device copy).
    DT *elements() const { return (DT *)lock(lm_read_and_write, tags::tv_cpu); }
    DT &operator[](unsigned long i) const { return elements()[i]; }

    DenseVector copy() const {
        DenseVector r(size());
        if (!_imp->host_valid && _imp->dev_valid) {
            void *d = r.lock(lm_write_only, tags::tv_gpu_hip);
            hip_check(honei_hip_copy(d, _imp->dev, sizeof(DT) * size(), nullptr), "DenseVector copy");
        } else if (size()) {
            std::memcpy(r._imp->host.data(), _imp->host.data(), sizeof(DT) * size());
        }
        return r;
    }

    bool same_storage(const DenseVector &o) const { return _imp == o._imp; }
    bool device_valid() const { return _imp->dev_valid; }
};

}  // namespace honei
