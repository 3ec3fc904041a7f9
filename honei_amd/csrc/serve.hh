// serve.hh -- continuous-batching request loop over a BatchMG pool.
//
// Requests (Q1 Poisson systems with a per-cell lognormal coefficient and a
// uniform right-hand side, drawn on the device from per-request seeds)
// arrive on a wall-clock schedule.  The pool's slots are sized at start-up
// from a device-memory budget; a request is admitted as soon as a slot of
// its size class is free, its hierarchy is assembled on the load stream
// while the other slots keep iterating, and it joins the next graph launch.
// Retired slots are reported at the poll after the launch that retired them
// and immediately reused.  Latency = completion poll time - arrival time.
#pragma once

#include <cstdint>
#include <vector>

#include "batch.hh"

namespace honei {
namespace hipx {

struct ServeRequest {
    double arrival_s;      // seconds after the start of the run
    int32_t level;         // side 2^level + 1
    uint64_t coeff_seed, rhs_seed;
    double *x_dev;         // optional: FP64 solution copied here on completion
    // optional host payload (pinned memory for asynchronous copies): the
    // coefficient ((side-1)^2) and right-hand side (side^2) are copied to the
    // device when the request is admitted (instead of being drawn from the
    // seeds), and the solution (side^2) is copied back before the request
    // counts as done
    const double *coeff_host = nullptr, *b_host = nullptr;
    double *x_host = nullptr;
};

struct ServeRecord {
    double admit_s = -1.0, done_s = -1.0, rel = 0.0;
    int32_t iters = 0, status = SLOT_PENDING;
    int64_t slot = -1;
};

struct ServeStats {
    int64_t completed = 0, not_converged = 0, rejected = 0, launches = 0, max_queue = 0;
    double wall_s = 0.0, solves_per_s = 0.0, p50_ms = 0.0, p99_ms = 0.0, max_ms = 0.0, mean_ms = 0.0;
    uint64_t device_bytes = 0;
};

struct ServeConfig {
    BatchConfig batch;
    int32_t nclass = 0;
    int64_t levels[kMaxClasses] = {};
    int64_t counts[kMaxClasses] = {};   // 0: sized from the budget
    uint64_t budget_bytes = 0;          // device bytes for the whole pool
    double sigma = 1.0;                 // lognormal coefficient parameter
    int32_t max_slots_per_class = 256;
};

// Device bytes one slot of finest level `level` occupies (threads-per-row 1
// layout; tuned layouts may need up to a third more matrix storage), plus the
// payload staging (coefficient and right-hand side) when `payload`.
uint64_t slot_bytes(int64_t level, ValueType inner, int outer_method, bool payload = false);

hipError_t serve_run(const ServeConfig &cfg, const ServeRequest *reqs, int64_t n, ServeRecord *out,
                     ServeStats *stats, bool *over_budget);

}  // namespace hipx
}  // namespace honei
