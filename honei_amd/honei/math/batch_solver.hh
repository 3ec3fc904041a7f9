// honei/math/batch_solver.hh -- many Q1 Poisson systems solved together on
// tags::GPU::HIP, and the continuous-batching request loop.
//
// No reference counterpart: the reference's Multigrid<..., MIXED>::value
// (math/multigrid.hh:2395-2536) solves one system per call.  BatchSolver
// owns a device-resident pool of slots (honei_hip_batch_*): each slot holds
// a rediscretised Q1 hierarchy built on the device from a per-cell
// coefficient, and all active slots run the MIXED defect correction together
// in graph launches.  ServeLoop runs the request loop (honei_hip_serve_run).
#pragma once

#include <cstdint>
#include <vector>

#include "honei/backends/hip/device.hh"

namespace honei {

struct BatchConfig : honei_hip_batch_config {
    BatchConfig() { hip_check(honei_hip_batch_config_default(this), "batch_config_default"); }
};

typedef honei_hip_slot_info SlotInfo;

template <typename Tag_> class BatchSolver;

template <> class BatchSolver<tags::GPU::HIP> {
    void *_h = nullptr;
    int64_t _nslots = 0;
    // inputs of loads that may still be in flight on the device
    std::vector<DenseVector<double>> _pending;

  public:
    // levels[c] (side 2^level + 1) with counts[c] slots each; budget 0 = none.
    BatchSolver(const std::vector<unsigned long> &levels, const std::vector<unsigned long> &counts,
                const BatchConfig &cfg = BatchConfig(), uint64_t budget_bytes = 0) {
        if (levels.size() != counts.size()) throw VectorSizeDoesNotMatch(counts.size(), levels.size());
        std::vector<int64_t> l(levels.begin(), levels.end()), c(counts.begin(), counts.end());
        hip_check(honei_hip_batch_create(&_h, l.data(), c.data(), (int32_t)l.size(), &cfg, budget_bytes),
                  "batch_create");
        hip_check(honei_hip_batch_info(_h, &_nslots, nullptr, nullptr), "batch_info");
    }
    ~BatchSolver() {
        if (_h) honei_hip_batch_destroy(_h);
    }
    BatchSolver(const BatchSolver &) = delete;
    BatchSolver &operator=(const BatchSolver &) = delete;

    unsigned long slots() const { return (unsigned long)_nslots; }

    uint64_t device_bytes() const {
        uint64_t b = 0;
        hip_check(honei_hip_batch_info(_h, nullptr, &b, nullptr), "batch_info");
        return b;
    }

    // Global slot of the j-th slot of class `cls`.
    unsigned long slot(unsigned cls, unsigned long j) const {
        int64_t g = 0;
        hip_check(honei_hip_batch_slot(_h, (int32_t)cls, (int64_t)j, &g), "batch_slot");
        return (unsigned long)g;
    }

    // coefficient: (side-1)^2 cell values; b: side^2 right-hand side.
    void load(unsigned long slot, const DenseVector<double> &coefficient, const DenseVector<double> &b) {
        hip_check(honei_hip_batch_load(_h, (int64_t)slot, (const double *)hip::rd(coefficient),
                                       (const double *)hip::rd(b)),
                  "batch_load");
        _pending.push_back(coefficient);
        _pending.push_back(b);
    }

    // One graph launch of every active slot (asynchronous).
    void step() { hip_check(honei_hip_batch_step(_h), "batch_step"); }

    // Wait for launched work; state of every slot.
    std::vector<SlotInfo> poll() {
        std::vector<SlotInfo> info((size_t)_nslots);
        hip_check(honei_hip_batch_poll(_h, info.data()), "batch_poll");
        return info;
    }

    // Launch until every loaded slot has retired; returns graph launches.
    unsigned long run() {
        int64_t k = 0;
        hip_check(honei_hip_batch_run(_h, &k), "batch_run");
        _pending.clear();
        return (unsigned long)k;
    }

    // FP64 solution of a retired slot.
    void result(unsigned long slot, DenseVector<double> &x) {
        hip_check(honei_hip_batch_result(_h, (int64_t)slot, (double *)hip::wr(x)), "batch_result");
    }
};

struct ServeLoop {
    // Runs the request loop to completion (requests sorted by arrival time).
    static honei_hip_serve_stats run(const honei_hip_serve_config &cfg, const std::vector<honei_hip_request> &reqs,
                                     std::vector<honei_hip_request_result> &results) {
        results.resize(reqs.size());
        honei_hip_serve_stats st{};
        hip_check(honei_hip_serve_run(&cfg, reqs.data(), (int64_t)reqs.size(), results.data(), &st), "serve_run");
        return st;
    }
};

}  // namespace honei
