// batch_capi.cpp -- extern "C" entries of the batched multigrid solver and
// the seeded field generators (include/honei_hip.h).

#include <hip/hip_runtime.h>

#include <new>
#include <string>
#include <vector>

#include "batch.hh"
#include "honei_hip.h"
#include "runtime.hh"
#include "serve.hh"

using namespace honei::hipx;

namespace {

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

BatchConfig to_config(const honei_hip_batch_config &c) {
    BatchConfig b;
    b.inner_vtype = c.inner_vtype;
    b.n_pre = c.n_pre;
    b.n_post = c.n_post;
    b.omega = c.omega;
    b.tolerance = c.tolerance;
    b.max_outer = c.max_outer;
    b.averaging = c.averaging;
    b.iters_per_launch = c.iters_per_launch;
    b.autotune = c.autotune;
    b.outer = c.outer_method;
    b.fused_levels = c.fused_levels;
    b.use_graphs = c.use_graphs;
    b.pair_sweeps = c.pair_sweeps;
    b.matrix_free = c.matrix_free;
    b.first_launch_from_history = c.first_launch_from_history;
    return b;
}

}  // namespace

extern "C" {

int honei_hip_fill_uniform(double *dst, size_t n, double lo, double hi, uint64_t seed, void *stream) {
    if (!dst && n) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null destination");
    return check(fill_uniform(dst, (int64_t)n, lo, hi, seed, (hipStream_t)stream), "fill_uniform");
}

int honei_hip_fill_lognormal(double *dst, size_t n, double sigma, uint64_t seed, void *stream) {
    if (!dst && n) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null destination");
    return check(fill_lognormal(dst, (int64_t)n, sigma, seed, (hipStream_t)stream), "fill_lognormal");
}

int honei_hip_batch_config_default(honei_hip_batch_config *cfg) {
    if (!cfg) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "cfg is null");
    const BatchConfig b;
    cfg->inner_vtype = b.inner_vtype;
    cfg->n_pre = b.n_pre;
    cfg->n_post = b.n_post;
    cfg->omega = b.omega;
    cfg->tolerance = b.tolerance;
    cfg->max_outer = b.max_outer;
    cfg->averaging = b.averaging;
    cfg->iters_per_launch = b.iters_per_launch;
    cfg->autotune = b.autotune;
    cfg->outer_method = b.outer;
    cfg->fused_levels = b.fused_levels;
    cfg->use_graphs = b.use_graphs;
    cfg->pair_sweeps = b.pair_sweeps;
    cfg->matrix_free = b.matrix_free;
    cfg->first_launch_from_history = b.first_launch_from_history;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_batch_create(void **batch, const int64_t *levels, const int64_t *counts, int32_t nclass,
                           const honei_hip_batch_config *cfg, uint64_t budget_bytes) {
    if (!batch || !levels || !counts || !cfg) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    *batch = nullptr;
    if (nclass < 1 || nclass > kMaxClasses) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "1..16 classes");
    int64_t total = 0;
    for (int32_t c = 0; c < nclass; ++c) {
        if (levels[c] < 2 || levels[c] > 13) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "class level outside 2..13");
        if (counts[c] < 1) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "class count < 1");
        total += counts[c];
    }
    if (total > 65535) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "more than 65535 slots");
    if (cfg->inner_vtype != HONEI_HIP_F32 && cfg->inner_vtype != HONEI_HIP_BF16 && cfg->inner_vtype != HONEI_HIP_F64)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "inner type must be F32, BF16 or F64");
    if (cfg->n_pre < 2 || cfg->n_pre > 16 || cfg->n_post < 0 || cfg->n_post > 16)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "sweeps: n_pre 2..16, n_post 0..16");
    if (cfg->iters_per_launch < 1 || cfg->iters_per_launch > 64 || cfg->max_outer < 1)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "iters_per_launch 1..64, max_outer >= 1");
    if (cfg->averaging < 0 || cfg->averaging > 2) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "averaging 0..2");
    if (cfg->outer_method != 0 && cfg->outer_method != 1)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "outer_method 0 or 1");
    if (cfg->fused_levels < 0 || cfg->fused_levels > kMaxFusedLevels)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "fused_levels 0..6");
    if (cfg->use_graphs != 0 && cfg->use_graphs != 1)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "use_graphs 0 or 1");
    if (cfg->pair_sweeps != 0 && cfg->pair_sweeps != 1)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "pair_sweeps 0 or 1");
    if (cfg->matrix_free < 0 || cfg->matrix_free > 2)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "matrix_free 0, 1 or 2");
    if (cfg->first_launch_from_history != 0 && cfg->first_launch_from_history != 1)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "first_launch_from_history 0 or 1");
    if (!(cfg->tolerance > 0.0) || !(cfg->omega > 0.0)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "tolerance, omega > 0");
    BatchMG *b = new (std::nothrow) BatchMG();
    if (!b) return fail(HONEI_HIP_ERROR_OUT_OF_MEMORY, "host allocation failed");
    const hipError_t e = b->create(levels, counts, nclass, to_config(*cfg), budget_bytes);
    if (e != hipSuccess) {
        const bool budget = b->over_budget();
        const uint64_t need = b->plan_bytes();
        delete b;
        if (budget)
            return fail(HONEI_HIP_ERROR_BUDGET, "batch pool needs " + std::to_string(need) + " bytes, budget " +
                                                    std::to_string(budget_bytes));
        return check(e, "batch_create");
    }
    *batch = b;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_batch_destroy(void *batch) {
    delete (BatchMG *)batch;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_batch_info(void *batch, int64_t *nslots, uint64_t *device_bytes, int64_t *launches) {
    if (!batch) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "batch is null");
    BatchMG *b = (BatchMG *)batch;
    if (nslots) *nslots = b->nslots();
    if (device_bytes) *device_bytes = b->bytes();
    if (launches) *launches = b->launches();
    return HONEI_HIP_SUCCESS;
}

int honei_hip_batch_slot(void *batch, int32_t cls, int64_t j, int64_t *slot) {
    if (!batch || !slot) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    const int64_t g = ((BatchMG *)batch)->slot(cls, j);
    if (g < 0) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "class/slot index out of range");
    *slot = g;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_batch_load(void *batch, int64_t slot, const double *coeff_dev, const double *b_dev) {
    if (!batch || !coeff_dev || !b_dev) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    BatchMG *b = (BatchMG *)batch;
    if (slot < 0 || slot >= b->nslots()) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "slot out of range");
    if (b->busy(slot)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "slot is loading or running");
    return check(b->load(slot, coeff_dev, b_dev), "batch_load");
}

static int check_slots(BatchMG *b, int64_t n, const int64_t *slots) {
    if (n < 0 || n > 65535 || (n && !slots)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad slot list");
    for (int64_t e = 0; e < n; ++e) {
        if (slots[e] < 0 || slots[e] >= b->nslots()) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "slot out of range");
        if (b->busy(slots[e])) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "slot is loading or running");
        for (int64_t f = 0; f < e; ++f)
            if (slots[f] == slots[e]) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "slot listed twice");
    }
    return HONEI_HIP_SUCCESS;
}

int honei_hip_batch_load_many(void *batch, int64_t n, const int64_t *slots, const double *const *coeff_dev,
                              const double *const *b_dev) {
    if (!batch || (n > 0 && (!coeff_dev || !b_dev))) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    BatchMG *b = (BatchMG *)batch;
    if (int st = check_slots(b, n, slots)) return st;
    for (int64_t e = 0; e < n; ++e)
        if (!coeff_dev[e] || !b_dev[e]) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null coefficient or rhs");
    return check(b->load_many(n, slots, coeff_dev, b_dev, nullptr, nullptr, 0.0), "batch_load_many");
}

int honei_hip_batch_load_seeded(void *batch, int64_t n, const int64_t *slots, const uint64_t *coeff_seeds,
                                const uint64_t *rhs_seeds, double sigma) {
    if (!batch || (n > 0 && (!coeff_seeds || !rhs_seeds))) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    BatchMG *b = (BatchMG *)batch;
    if (int st = check_slots(b, n, slots)) return st;
    return check(b->load_many(n, slots, nullptr, nullptr, coeff_seeds, rhs_seeds, sigma), "batch_load_seeded");
}

int honei_hip_batch_activate(void *batch, int32_t wait) {
    if (!batch) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "batch is null");
    return check(((BatchMG *)batch)->activate_ready(wait != 0), "batch_activate");
}

int honei_hip_batch_step(void *batch) {
    if (!batch) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "batch is null");
    return check(((BatchMG *)batch)->step(), "batch_step");
}

int honei_hip_batch_poll(void *batch, honei_hip_slot_info *info) {
    if (!batch) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "batch is null");
    BatchMG *b = (BatchMG *)batch;
    std::vector<SlotScalars> st;
    const int rc = check(b->poll(st), "batch_poll");
    if (rc || !info) return rc;
    for (int64_t g = 0; g < b->nslots(); ++g) {
        const SlotScalars &s = st[(size_t)g];
        info[g].initial = s.initial;
        info[g].norm = s.norm;
        info[g].rel = s.rel;
        info[g].state = b->loading(g) ? 3 : (s.state == SLOT_VERIFY ? SLOT_ACTIVE : s.state);
        info[g].iters = s.iters;
        info[g].status = s.status;
        info[g].level = (int32_t)b->slot_level(g);
    }
    return HONEI_HIP_SUCCESS;
}

int honei_hip_batch_result(void *batch, int64_t slot, double *x_dev) {
    if (!batch || !x_dev) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    BatchMG *b = (BatchMG *)batch;
    if (slot < 0 || slot >= b->nslots()) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "slot out of range");
    return check(b->result(slot, x_dev), "batch_result");
}

int honei_hip_batch_run(void *batch, int64_t *launches) {
    if (!batch) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "batch is null");
    return check(((BatchMG *)batch)->run_to_completion(launches), "batch_run");
}

int honei_hip_batch_shape(void *batch, int32_t level, int32_t *threads, int32_t *block, int32_t *group,
                          double *sweep_us, int64_t *tune_slots) {
    if (!batch) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "batch is null");
    BatchMG *b = (BatchMG *)batch;
    const auto &L = b->levels();
    if (level < 1 || level >= (int32_t)L.size()) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "level out of range");
    if (threads) *threads = L[(size_t)level].T;
    if (block) *block = L[(size_t)level].block;
    if (group) *group = L[(size_t)level].group;
    LaunchShape sh;
    int dev = 0;
    const bool tuned = hipGetDevice(&dev) == hipSuccess &&
                       batch_tuned_shape(dev, level, b->config().inner_vtype, &sh) && b->config().autotune;
    if (sweep_us) *sweep_us = tuned ? sh.us : 0.0;
    if (tune_slots) *tune_slots = tuned ? sh.slots : 0;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_batch_pair_timing(void *batch, int32_t level, double *pre_us, double *post_us) {
    if (!batch) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "batch is null");
    BatchMG *b = (BatchMG *)batch;
    const auto &L = b->levels();
    if (level < 1 || level >= (int32_t)L.size()) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "level out of range");
    LaunchShape sh;
    int dev = 0;
    const bool tuned = hipGetDevice(&dev) == hipSuccess &&
                       batch_tuned_shape(dev, level, b->config().inner_vtype, &sh) && b->config().autotune &&
                       b->paired(L[(size_t)level]);
    const int mf = b->config().matrix_free;
    if (pre_us) *pre_us = tuned ? sh.pre_us[mf] : 0.0;
    if (post_us) *post_us = tuned ? sh.post_us[mf] : 0.0;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_batch_time_level(void *batch, int32_t level, int32_t reps, double *us) {
    if (!batch || !us) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    BatchMG *b = (BatchMG *)batch;
    if (level < 2 || level >= (int32_t)b->levels().size()) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "level out of range");
    if (reps < 1 || reps > 1000) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "reps outside 1..1000");
    if (b->loading_count() > 0 || b->running() > 0)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "the pool has slots loading or running");
    return check(b->time_level(level, reps, us), "batch_time_level");
}

int honei_hip_serve_run(const honei_hip_serve_config *cfg, const honei_hip_request *reqs, int64_t nreq,
                        honei_hip_request_result *results, honei_hip_serve_stats *stats) {
    if (!cfg || (nreq > 0 && (!reqs || !results)) || nreq < 0)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    if (cfg->nclass < 1 || cfg->nclass > kMaxClasses) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "1..16 classes");
    ServeConfig sc;
    sc.batch = to_config(cfg->batch);
    sc.nclass = cfg->nclass;
    for (int c = 0; c < cfg->nclass; ++c) {
        if (cfg->levels[c] < 2 || cfg->levels[c] > 13) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "class level outside 2..13");
        if (cfg->counts[c] < 0) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "negative class count");
        if (cfg->counts[c] == 0 && cfg->budget_bytes == 0)
            return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "class sized from the budget needs budget_bytes");
        sc.levels[c] = cfg->levels[c];
        sc.counts[c] = cfg->counts[c];
    }
    sc.budget_bytes = cfg->budget_bytes;
    sc.sigma = cfg->sigma;
    sc.max_slots_per_class = cfg->max_slots_per_class > 0 ? cfg->max_slots_per_class : 256;
    std::vector<ServeRequest> rq((size_t)nreq);
    for (int64_t r = 0; r < nreq; ++r) {
        if (r > 0 && reqs[r].arrival_s < reqs[r - 1].arrival_s)
            return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "arrival times must be non-decreasing");
        if ((reqs[r].coeff_host == nullptr) != (reqs[r].b_host == nullptr))
            return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "coeff_host and b_host: both or neither");
        rq[(size_t)r] = ServeRequest{reqs[r].arrival_s, reqs[r].level, reqs[r].coeff_seed, reqs[r].rhs_seed,
                                     reqs[r].x_dev, reqs[r].coeff_host, reqs[r].b_host, reqs[r].x_host};
    }
    std::vector<ServeRecord> rec((size_t)nreq);
    ServeStats st;
    bool budget = false;
    const hipError_t e = serve_run(sc, rq.data(), nreq, rec.data(), &st, &budget);
    if (budget) return fail(HONEI_HIP_ERROR_BUDGET, "serving pool exceeds the memory budget");
    if (e != hipSuccess) return check(e, "serve_run");
    for (int64_t r = 0; r < nreq; ++r) {
        const ServeRecord &x = rec[(size_t)r];
        results[r].admit_s = x.admit_s;
        results[r].done_s = x.done_s;
        results[r].rel = x.rel;
        results[r].iters = x.iters;
        results[r].status = x.status;
        results[r].slot = x.slot;
    }
    if (stats) {
        stats->completed = st.completed;
        stats->not_converged = st.not_converged;
        stats->rejected = st.rejected;
        stats->launches = st.launches;
        stats->max_queue = st.max_queue;
        stats->wall_s = st.wall_s;
        stats->solves_per_s = st.solves_per_s;
        stats->p50_ms = st.p50_ms;
        stats->p99_ms = st.p99_ms;
        stats->max_ms = st.max_ms;
        stats->mean_ms = st.mean_ms;
        stats->device_bytes = st.device_bytes;
    }
    return HONEI_HIP_SUCCESS;
}

}  // extern "C"
