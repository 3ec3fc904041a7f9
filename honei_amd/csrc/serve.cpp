// serve.cpp -- continuous-batching request loop (see serve.hh).

#include "serve.hh"

#include <algorithm>
#include <cmath>
#include <chrono>
#include <deque>
#include <thread>

namespace honei {
namespace hipx {

uint64_t slot_bytes(int64_t level, ValueType inner, int outer_method, bool payload) {
    const uint64_t vs = vt_size(inner), es = inner == VT_F64 ? 8 : 4;   // value, inner vector element
    uint64_t b = 0;
    for (int64_t l = 1; l <= level; ++l) {
        const uint64_t side = (1ull << l) + 1, n = side * side;
        const uint64_t stride = 32 * ((n + 31) / 32), vld = (n + 63) / 64 * 64;
        const uint64_t kld = ((side - 1) * (side - 1) + 31) / 32 * 32;
        b += 9 * stride * vs + vld * es * (l == 1 ? 8 : 5) + kld * 8;   // values, vectors, cell coefficients
    }
    const uint64_t side = (1ull << level) + 1, n = side * side;
    const uint64_t stride = 32 * ((n + 31) / 32), vld = (n + 63) / 64 * 64;
    b += 9 * stride * 8 + vld * 8 * (outer_method == 1 ? 6 : 3) + 4 * ((n + 255) / 256) * 8 + sizeof(SlotScalars) + 8;
    if (payload) b += ((side - 1) * (side - 1) + n) * 8;
    return b;
}

namespace {

// Per-slot device staging of host payloads (coefficient, right-hand side).
struct Staging {
    std::vector<double *> coeff, rhs;
    std::vector<void *> allocs;
    ~Staging() {
        for (void *p : allocs) (void)hipFree(p);
    }
};

}  // namespace

hipError_t serve_run(const ServeConfig &cfg, const ServeRequest *reqs, int64_t n, ServeRecord *out,
                     ServeStats *stats, bool *over_budget) {
    using clock = std::chrono::steady_clock;
    if (over_budget) *over_budget = false;
    if (cfg.nclass < 1 || cfg.nclass > kMaxClasses || n < 0) return hipErrorInvalidValue;
    bool payload = false;
    for (int64_t r = 0; r < n; ++r) {
        if ((reqs[r].coeff_host == nullptr) != (reqs[r].b_host == nullptr)) return hipErrorInvalidValue;
        payload = payload || reqs[r].coeff_host;
    }
    // pool: explicit class sizes, or an equal share of the budget per class
    // (matrix storage counted at the largest tuned layout, 12/9 of T = 1)
    int64_t counts[kMaxClasses];
    for (int c = 0; c < cfg.nclass; ++c) {
        counts[c] = cfg.counts[c];
        if (counts[c] > 0) continue;
        if (!cfg.budget_bytes) return hipErrorInvalidValue;
        const uint64_t per =
            slot_bytes(cfg.levels[c], (ValueType)cfg.batch.inner_vtype, cfg.batch.outer, payload) * 4 / 3 + 1;
        const uint64_t share = cfg.budget_bytes / (uint64_t)cfg.nclass;
        counts[c] = std::max<int64_t>(1, std::min<int64_t>(cfg.max_slots_per_class, (int64_t)(share / per)));
    }
    Staging stage;   // declared first: released after the pool has drained its streams
    BatchMG pool;
    const hipError_t ce = pool.create(cfg.levels, counts, cfg.nclass, cfg.batch, cfg.budget_bytes);
    if (ce != hipSuccess) {
        if (over_budget) *over_budget = pool.over_budget();
        return ce;
    }
    hipStream_t s_ret = nullptr;
    if (payload) {
        stage.coeff.assign((size_t)pool.nslots(), nullptr);
        stage.rhs.assign((size_t)pool.nslots(), nullptr);
        for (int c = 0; c < cfg.nclass; ++c) {
            const int64_t side = (1ll << cfg.levels[c]) + 1, cells = (side - 1) * (side - 1), nn = side * side;
            double *kc = nullptr, *bc = nullptr;
            HX(hipMalloc((void **)&kc, (size_t)(counts[c] * cells) * 8));
            stage.allocs.push_back(kc);
            HX(hipMalloc((void **)&bc, (size_t)(counts[c] * nn) * 8));
            stage.allocs.push_back(bc);
            for (int64_t j = 0; j < counts[c]; ++j) {
                const int64_t g = pool.slot(c, j);
                stage.coeff[(size_t)g] = kc + j * cells;
                stage.rhs[(size_t)g] = bc + j * nn;
            }
        }
        HX(hipStreamCreateWithFlags(&s_ret, hipStreamNonBlocking));
    }
    std::vector<std::deque<int64_t>> queue((size_t)cfg.nclass);
    std::vector<std::vector<int64_t>> free_slots((size_t)cfg.nclass);
    for (int c = 0; c < cfg.nclass; ++c)
        for (int64_t j = counts[c] - 1; j >= 0; --j) free_slots[(size_t)c].push_back(pool.slot(c, j));
    std::vector<int64_t> slot_req((size_t)pool.nslots(), -1);
    std::vector<int> slot_class((size_t)pool.nslots(), 0);
    for (int c = 0; c < cfg.nclass; ++c)
        for (int64_t j = 0; j < counts[c]; ++j) slot_class[(size_t)pool.slot(c, j)] = c;
    auto class_of = [&](int32_t level) {
        for (int c = 0; c < cfg.nclass; ++c)
            if (cfg.levels[c] == level) return c;
        return -1;
    };
    for (int64_t r = 0; r < n; ++r) out[r] = ServeRecord();

    ServeStats st;
    st.device_bytes = pool.bytes();
    std::vector<SlotScalars> sc;
    std::vector<int64_t> retired, adm_slots, pay_slots;
    std::vector<uint64_t> adm_cs, adm_bs;
    std::vector<const double *> pay_k, pay_b;
    // solutions on their way back to host memory: (slot, completion event)
    std::vector<std::pair<int64_t, hipEvent_t>> returning;
    std::vector<hipEvent_t> spare_ev;
    auto take_event = [&](hipEvent_t *e) -> hipError_t {
        if (!spare_ev.empty()) {
            *e = spare_ev.back();
            spare_ev.pop_back();
            return hipSuccess;
        }
        return hipEventCreateWithFlags(e, hipEventDisableTiming);
    };
    struct EventPool {
        std::vector<hipEvent_t> *a;
        std::vector<std::pair<int64_t, hipEvent_t>> *b;
        hipStream_t s;
        ~EventPool() {
            if (s) (void)hipStreamSynchronize(s);
            for (hipEvent_t e : *a) (void)hipEventDestroy(e);
            for (auto &p : *b) (void)hipEventDestroy(p.second);
            if (s) (void)hipStreamDestroy(s);
        }
    } event_guard{&spare_ev, &returning, s_ret};
    const auto t0 = clock::now();
    auto elapsed = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
    int64_t next = 0, finished = 0, queued = 0;
    auto finish = [&](int64_t g, const SlotScalars &s, double t) {
        const int64_t r = slot_req[(size_t)g];
        out[r].done_s = t;
        out[r].iters = s.iters;
        out[r].status = s.status;
        out[r].rel = s.rel;
        if (s.status == SLOT_OK) ++st.completed;
        else ++st.not_converged;
        slot_req[(size_t)g] = -1;
        free_slots[(size_t)slot_class[(size_t)g]].push_back(g);
        ++finished;
    };
    std::vector<SlotScalars> ret_state((size_t)pool.nslots());
    while (finished < n) {
        const double now = elapsed();
        for (; next < n && reqs[next].arrival_s <= now; ++next) {
            const int c = class_of(reqs[next].level);
            if (c < 0) {
                out[next].status = -2;   // no slot class for this size
                ++st.rejected;
                ++finished;
                continue;
            }
            queue[(size_t)c].push_back(next);
            ++queued;
        }
        st.max_queue = std::max(st.max_queue, queued);
        // admit into free slots, one bulk build for everything admitted now
        // (one for seeded requests, one for host payloads staged first)
        adm_slots.clear();
        adm_cs.clear();
        adm_bs.clear();
        pay_slots.clear();
        pay_k.clear();
        pay_b.clear();
        for (int c = 0; c < cfg.nclass; ++c) {
            auto &q = queue[(size_t)c];
            auto &fs = free_slots[(size_t)c];
            while (!q.empty() && !fs.empty()) {
                const int64_t r = q.front(), g = fs.back();
                q.pop_front();
                fs.pop_back();
                --queued;
                if (reqs[r].coeff_host) {
                    const int64_t side = (1ll << reqs[r].level) + 1;
                    HX(hipMemcpyAsync(stage.coeff[(size_t)g], reqs[r].coeff_host, (size_t)((side - 1) * (side - 1)) * 8,
                                      hipMemcpyHostToDevice, pool.load_stream()));
                    HX(hipMemcpyAsync(stage.rhs[(size_t)g], reqs[r].b_host, (size_t)(side * side) * 8,
                                      hipMemcpyHostToDevice, pool.load_stream()));
                    pay_slots.push_back(g);
                    pay_k.push_back(stage.coeff[(size_t)g]);
                    pay_b.push_back(stage.rhs[(size_t)g]);
                } else {
                    adm_slots.push_back(g);
                    adm_cs.push_back(reqs[r].coeff_seed);
                    adm_bs.push_back(reqs[r].rhs_seed);
                }
                slot_req[(size_t)g] = r;
                out[r].slot = g;
                out[r].admit_s = elapsed();
            }
        }
        if (!adm_slots.empty())
            HX(pool.load_many((int64_t)adm_slots.size(), adm_slots.data(), nullptr, nullptr, adm_cs.data(),
                              adm_bs.data(), cfg.sigma));
        if (!pay_slots.empty())
            HX(pool.load_many((int64_t)pay_slots.size(), pay_slots.data(), pay_k.data(), pay_b.data(), nullptr,
                              nullptr, cfg.sigma));
        // solutions whose copy to host memory has completed
        for (size_t k = 0; k < returning.size();) {
            const hipError_t q = hipEventQuery(returning[k].second);
            if (q == hipErrorNotReady) {
                ++k;
                continue;
            }
            HX(q);
            const int64_t g = returning[k].first;
            finish(g, ret_state[(size_t)g], elapsed());
            spare_ev.push_back(returning[k].second);
            returning[k] = returning.back();
            returning.pop_back();
        }
        if (pool.running() + pool.loading_count() == 0) {
            if (!returning.empty()) {
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            } else if (next < n) {
                const double wait = reqs[next].arrival_s - elapsed();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.002)));
            }
            continue;
        }
        if (pool.running() == 0) HX(pool.activate_ready(true));   // nothing to overlap with
        HX(pool.step());
        HX(pool.poll(sc, &retired));
        const double t = elapsed();
        for (int64_t g : retired) {
            const int64_t r = slot_req[(size_t)g];
            if (r < 0) continue;
            if (reqs[r].x_dev) HX(pool.result(g, reqs[r].x_dev));
            if (reqs[r].x_host) {   // the slot stays taken until its solution is in host memory
                int64_t nn = 0;
                const double *xs = pool.solution(g, &nn);
                hipEvent_t e = nullptr;
                HX(take_event(&e));
                HX(hipMemcpyAsync(reqs[r].x_host, xs, (size_t)nn * 8, hipMemcpyDeviceToHost, s_ret));
                HX(hipEventRecord(e, s_ret));
                ret_state[(size_t)g] = sc[(size_t)g];
                returning.push_back({g, e});
                continue;
            }
            finish(g, sc[(size_t)g], t);
        }
    }
    st.wall_s = elapsed();
    st.launches = pool.launches();
    std::vector<double> lat;
    for (int64_t r = 0; r < n; ++r)
        if (out[r].status == SLOT_OK) lat.push_back(1000.0 * (out[r].done_s - reqs[r].arrival_s));
    if (!lat.empty()) {
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
            const size_t k = (size_t)std::min<double>((double)lat.size() - 1, std::ceil(p * (double)lat.size()) - 1);
            return lat[k];
        };
        st.p50_ms = pct(0.50);
        st.p99_ms = pct(0.99);
        st.max_ms = lat.back();
        double sum = 0.0;
        for (double v : lat) sum += v;
        st.mean_ms = sum / (double)lat.size();
    }
    st.solves_per_s = st.wall_s > 0 ? (double)st.completed / st.wall_s : 0.0;
    if (stats) *stats = st;
    return hipSuccess;
}

}  // namespace hipx
}  // namespace honei
