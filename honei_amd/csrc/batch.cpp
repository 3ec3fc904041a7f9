// batch.cpp -- host side of the batched multigrid solver: pool layout,
// launch-shape autotuning, per-slot loading, graph capture and the outer
// iteration loop.

#include "batch.hh"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

namespace honei {
namespace hipx {

namespace {

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// Tiles of k_b_dir_spmv over a side x side grid: the paired passes' column
// strips (at most kPairWidth columns) in bands of kDirRows rows.
struct DirTiles {
    int32_t tw, strips;
    int64_t ntiles;
};
DirTiles dir_tiles(int64_t side) {
    DirTiles t;
    const int64_t strips = (side + kPairWidth - 1) / kPairWidth;
    t.strips = (int32_t)strips;
    t.tw = (int32_t)((side + strips - 1) / strips);
    t.ntiles = strips * ((side + kDirRows - 1) / kDirRows);
    return t;
}

// Launch shapes chosen by the autotuner, per (device, level, value type), for
// the lifetime of the process: every batch of a process uses one shape per
// level, so a system's result does not depend on the other systems in its
// batch.
std::mutex g_tune_mu;
std::map<std::tuple<int, int64_t, int>, LaunchShape> g_tuned;

struct Level0 {
    int64_t side, n, stride, ncpr, ax_ld, vld, bps;
};

Level0 level_geometry(int64_t level, int T, int block) {
    Level0 g;
    g.side = (1ll << level) + 1;
    g.n = g.side * g.side;
    g.stride = 32 * ((g.n * T + 31) / 32);
    g.ncpr = (9 + T - 1) / T;
    g.ax_ld = g.ncpr * g.stride;
    g.vld = round_up(g.n, 64);
    g.bps = (g.n * T + block - 1) / block;
    return g;
}

}  // namespace

BatchMG::~BatchMG() { release(); }

// step() and load_many() return before their work has run, so everything
// the launched graphs and builds read (graph executables with their kernel
// arguments, pool memory, staging buffers) is released only after both
// streams have drained.
void BatchMG::release() {
    if (s_main_) (void)hipStreamSynchronize(s_main_);
    if (s_load_) (void)hipStreamSynchronize(s_load_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (exec1_) (void)hipGraphExecDestroy(exec1_);
    if (graph1_) (void)hipGraphDestroy(graph1_);
    if (execf_) (void)hipGraphExecDestroy(execf_);
    if (graphf_) (void)hipGraphDestroy(graphf_);
    exec_ = exec1_ = execf_ = nullptr;
    graph_ = graph1_ = graphf_ = nullptr;
    for (hipEvent_t e : load_ev_)
        if (e) (void)hipEventDestroy(e);
    load_ev_.clear();
    if (null_ev_) (void)hipEventDestroy(null_ev_);
    null_ev_ = nullptr;
    for (void *p : allocs_) (void)hipFree(p);
    allocs_.clear();
    if (sc_host_) (void)hipHostFree(sc_host_);
    sc_host_ = nullptr;
    if (s_main_) (void)hipStreamDestroy(s_main_);
    if (s_load_) (void)hipStreamDestroy(s_load_);
    s_main_ = s_load_ = nullptr;
    if (build_ev_) (void)hipEventDestroy(build_ev_);
    build_ev_ = nullptr;
    if (build_host_) (void)hipHostFree(build_host_);
    build_host_ = nullptr;
    if (build_dev_) (void)hipFree(build_dev_);
    build_dev_ = nullptr;
    build_cap_ = 0;
    bytes_ = 0;
}

hipError_t BatchMG::alloc(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    HX(hipMalloc(p, bytes));
    allocs_.push_back(*p);
    bytes_ += bytes;
    return fill_bytes_zero(*p, bytes, s_main_);
}

int64_t BatchMG::slot(int cls, int64_t j) const {
    if (cls < 0 || cls >= (int)class_of_.size()) return -1;
    const BClass &c = C_[(size_t)class_of_[(size_t)cls]];
    if (j < 0 || j >= c.count) return -1;
    return c.first + j;
}

int64_t BatchMG::slot_level(int64_t g) const {
    for (const BClass &c : C_)
        if (g >= c.first && g < c.first + c.count) return c.level;
    return -1;
}

// Time one batched Jacobi sweep for each (T, block) candidate on a synthetic
// batch of the level's shape (constant-coefficient values; timing does not
// depend on them) and keep the fastest.
hipError_t BatchMG::tune_level(BLevel &lv) {
    const ValueType vt = (ValueType)cfg_.inner_vtype;
    int dev = 0;
    HX(hipGetDevice(&dev));
    const auto key = std::make_tuple(dev, lv.level, (int)vt);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            lv.T = it->second.T;
            lv.block = it->second.block;
            lv.group = it->second.group;
            return hipSuccess;
        }
    }
    const int64_t n = lv.n;
    const int64_t ns = std::max<int64_t>(1, std::min<int64_t>(256, (int64_t(1) << 22) / n));
    LaunchShape best;
    best.us = 1e30;
    hipEvent_t e0, e1;
    HX(hipEventCreate(&e0));
    HX(hipEventCreate(&e1));
    hipError_t err = hipSuccess;
    const size_t vs = vt_size(vt), es = inner_vector_bytes();
    for (int T : {1, 2, 4}) {
        const Level0 g = level_geometry(lv.level, T, 256);
        void *Aj = nullptr, *Arl = nullptr, *Ax = nullptr, *vec = nullptr, *scd = nullptr, *lst = nullptr;
        const size_t vbytes = (size_t)(ns * g.vld) * es;
        err = hipMalloc(&Aj, (size_t)g.ax_ld * 4);
        if (err == hipSuccess) err = hipMalloc(&Arl, (size_t)n * 4);
        if (err == hipSuccess) err = hipMalloc(&Ax, (size_t)(ns * g.ax_ld) * vs);
        if (err == hipSuccess) err = hipMalloc(&vec, 4 * vbytes);
        if (err == hipSuccess) err = hipMalloc(&scd, (size_t)ns * sizeof(SlotScalars));
        if (err == hipSuccess) err = hipMalloc(&lst, (size_t)(2 * ns + 2) * sizeof(int32_t));
        if (err == hipSuccess) err = fill_bytes_zero(vec, 4 * vbytes, s_main_);
        if (err == hipSuccess)
            err = q1_assemble(vt, g.side, (int32_t *)Aj, Ax, (uint32_t *)Arl, g.stride, nullptr, vec, 0.8, s_main_, T);
        for (int64_t k = 1; k < ns && err == hipSuccess; ++k)
            err = hipMemcpyAsync((char *)Ax + (size_t)(k * g.ax_ld) * vs, Ax, (size_t)g.ax_ld * vs,
                                 hipMemcpyDeviceToDevice, s_main_);
        if (err == hipSuccess) err = b_fill_state((SlotScalars *)scd, ns, SLOT_ACTIVE, s_main_);
        if (err == hipSuccess) {
            const int64_t bnd[2] = {0, ns};
            err = b_compact((const SlotScalars *)scd, ns, bnd, 2, (int32_t *)lst, (int32_t *)lst + ns, LIST_NONIDLE,
                            s_main_);
        }
        const BWork w{(const int32_t *)lst, (const int32_t *)lst + ns, 0, 1, 0};
        for (int block : {64, 128, 256, 512, 1024}) {
            for (int group : {1, 2, 4}) {
                if (err != hipSuccess) break;
                if (block < T || group > ns) continue;
                BMat m{(const int32_t *)Aj, (const uint32_t *)Arl, Ax, g.stride, g.ax_ld, n, g.vld,
                       (n * T + block - 1) / block};
                char *v0 = (char *)vec;
                void *dinv = v0, *x = v0 + vbytes, *rhs = v0 + 2 * vbytes, *out = v0 + 3 * vbytes;
                const int reps = n >= 65536 ? 10 : 30;
                for (int r = 0; r < 3 && err == hipSuccess; ++r)
                    err = b_ell(vt, T, block, group, ELL_JACOBI, m, w, ns, out, x, rhs, dinv, s_main_);
                if (err == hipSuccess) err = hipEventRecord(e0, s_main_);
                for (int r = 0; r < reps && err == hipSuccess; ++r)
                    err = b_ell(vt, T, block, group, ELL_JACOBI, m, w, ns, out, x, rhs, dinv, s_main_);
                if (err == hipSuccess) err = hipEventRecord(e1, s_main_);
                if (err == hipSuccess) err = hipEventSynchronize(e1);
                float ms = 0.f;
                if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
                const double us = 1000.0 * ms / reps;
                if (err == hipSuccess && us < best.us) {
                    best.us = us;
                    best.T = T;
                    best.block = block;
                    best.group = group;
                }
            }
        }
        if (T == 1 && paired(lv) && err == hipSuccess) {
            // the paired launches this level will run, timed the same way
            BLevel tl;
            tl.level = lv.level;
            tl.side = g.side;
            tl.n = n;
            tl.cap = ns;
            tl.vld = g.vld;
            tl.T = 1;
            tl.stride = g.stride;
            tl.ncpr = g.ncpr;
            tl.ax_ld = g.ax_ld;
            tl.Aj = (int32_t *)Aj;
            tl.Arl = (uint32_t *)Arl;
            tl.Ax = Ax;
            char *v0 = (char *)vec;
            tl.dinv = v0;
            tl.x = v0 + vbytes;
            tl.rhs = v0 + 2 * vbytes;
            void *out = v0 + 3 * vbytes;
            tl.kld = round_up((g.side - 1) * (g.side - 1), 32);
            err = hipMalloc((void **)&tl.kappa, (size_t)(ns * tl.kld) * 8);
            if (err == hipSuccess) err = vfill(tl.kappa, 1.0, ns * tl.kld, VT_F64, s_main_);
            if (err == hipSuccess) err = hipMalloc(&tl.st, (size_t)(ns * 4 * g.vld) * vs);
            if (err == hipSuccess) err = b_stencil_fill(vt, g.side, tl.kappa, tl.kld, tl.st, g.vld, ns, s_main_);
            // every weight source: the stored values (0), the cell
            // coefficients (1), the distinct-weight arrays (2)
            for (int mf = 0; mf < 3; ++mf) {
                for (int pass = 0; pass < 2 && err == hipSuccess; ++pass) {
                    auto launch = [&]() {
                        return pass == 0 ? b_pair(vt, PAIR_PRE, tl, w, tl.x, out, nullptr, s_main_, mf)
                                         : b_pair(vt, PAIR_POST, tl, w, out, nullptr, tl.x, s_main_, mf);
                    };
                    const int reps = n >= 65536 ? 10 : 30;
                    for (int r = 0; r < 3 && err == hipSuccess; ++r) err = launch();
                    if (err == hipSuccess) err = hipEventRecord(e0, s_main_);
                    for (int r = 0; r < reps && err == hipSuccess; ++r) err = launch();
                    if (err == hipSuccess) err = hipEventRecord(e1, s_main_);
                    if (err == hipSuccess) err = hipEventSynchronize(e1);
                    float ms = 0.f;
                    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
                    (pass == 0 ? best.pre_us[mf] : best.post_us[mf]) = 1000.0 * ms / reps;
                }
            }
            (void)hipStreamSynchronize(s_main_);
            if (tl.kappa) (void)hipFree(tl.kappa);
            if (tl.st) (void)hipFree(tl.st);
        }
        (void)hipStreamSynchronize(s_main_);
        for (void *p : {Aj, Arl, Ax, vec, scd, lst})
            if (p) (void)hipFree(p);
        if (err != hipSuccess) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HX(err);
    lv.T = best.T;
    lv.block = best.block;
    lv.group = best.group;
    best.slots = ns;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tuned[key] = best;
    return hipSuccess;
}

bool batch_tuned_shape(int dev, int64_t level, int vt, LaunchShape *out) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple(dev, level, vt));
    if (it == g_tuned.end()) return false;
    *out = it->second;
    return true;
}

uint64_t BatchMG::plan_bytes() const {
    uint64_t b = 0;
    const size_t vs = vt_size((ValueType)cfg_.inner_vtype);
    for (int64_t l = 1; l <= max_level_; ++l) {
        const BLevel &lv = L_[(size_t)l];
        if (stores_values(lv)) {
            b += (uint64_t)lv.ax_ld * 4 + (uint64_t)lv.n * 4;                      // Aj, Arl
            b += (uint64_t)(lv.cap * lv.ax_ld) * vs;                                // Ax
        }
        b += (uint64_t)(lv.cap * lv.vld) * inner_vector_bytes() * (l == 1 ? 8 : 5);   // vectors
        b += (uint64_t)(lv.cap * lv.kld) * 8;                                       // cell coefficients
        if (stores_stencil(lv)) b += (uint64_t)(lv.cap * 4 * lv.vld) * vs;          // distinct weights
    }
    for (const BClass &c : C_) {
        if (stores_outer_values()) b += (uint64_t)c.ax_ld * 4 + (uint64_t)c.n * 4 + (uint64_t)(c.count * c.ax_ld) * 8;
        b += (uint64_t)(c.count * c.vld) * 8 * (cfg_.outer == 1 ? 6 : 3);
        b += (uint64_t)(c.count * 4 * c.bps) * 8;   // partials
    }
    b += (uint64_t)nslots_ * sizeof(SlotScalars);
    return b;
}

hipError_t BatchMG::create(const int64_t *levels, const int64_t *counts, int nclass, const BatchConfig &cfg,
                           uint64_t budget_bytes) {
    release();
    over_budget_ = false;
    cfg_ = cfg;
    if (nclass < 1 || nclass > kMaxClasses) return hipErrorInvalidValue;
    if (cfg.inner_vtype != VT_F32 && cfg.inner_vtype != VT_BF16 && cfg.inner_vtype != VT_F64)
        return hipErrorInvalidValue;
    if (cfg.n_pre < 2 || cfg.n_pre > 16 || cfg.n_post < 0 || cfg.n_post > 16) return hipErrorInvalidValue;
    if (cfg.iters_per_launch < 1 || cfg.iters_per_launch > 64 || cfg.max_outer < 1) return hipErrorInvalidValue;
    if (cfg.averaging < 0 || cfg.averaging > 2) return hipErrorInvalidValue;
    if (cfg.outer != 0 && cfg.outer != 1) return hipErrorInvalidValue;
    if (cfg.use_graphs != 0 && cfg.use_graphs != 1) return hipErrorInvalidValue;
    if (cfg.pair_sweeps != 0 && cfg.pair_sweeps != 1) return hipErrorInvalidValue;
    if (cfg.matrix_free < 0 || cfg.matrix_free > 2) return hipErrorInvalidValue;
    std::vector<int> order((size_t)nclass);
    for (int c = 0; c < nclass; ++c) {
        if (levels[c] < 2 || levels[c] > 13 || counts[c] < 1 || counts[c] > 65535) return hipErrorInvalidValue;
        order[(size_t)c] = c;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return levels[a] > levels[b]; });
    class_of_.assign((size_t)nclass, 0);
    C_.assign((size_t)nclass, BClass());
    int64_t first = 0;
    for (int k = 0; k < nclass; ++k) {
        const int c = order[(size_t)k];
        class_of_[(size_t)c] = k;
        BClass &bc = C_[(size_t)k];
        const Level0 g = level_geometry(levels[c], 1, 256);
        bc.level = levels[c];
        bc.side = g.side;
        bc.n = g.n;
        bc.first = first;
        bc.count = counts[c];
        bc.vld = g.vld;
        bc.stride = g.stride;
        bc.ax_ld = g.ax_ld;
        bc.bps = g.bps;
        first += counts[c];
    }
    nslots_ = first;
    if (nslots_ > 65535) return hipErrorInvalidValue;
    max_level_ = C_[0].level;

    HX(hipStreamCreateWithFlags(&s_main_, hipStreamNonBlocking));
    HX(hipStreamCreateWithFlags(&s_load_, hipStreamNonBlocking));
    HX(hipEventCreateWithFlags(&build_ev_, hipEventDisableTiming));
    HX(hipEventCreateWithFlags(&null_ev_, hipEventDisableTiming));

    L_.assign((size_t)max_level_ + 1, BLevel());
    for (int64_t l = 1; l <= max_level_; ++l) {
        BLevel &lv = L_[(size_t)l];
        lv.level = l;
        lv.side = (1ll << l) + 1;
        lv.n = lv.side * lv.side;
        for (const BClass &c : C_)
            if (c.level >= l) lv.cap += c.count;
        lv.T = 1;
        lv.block = 256;
        lv.group = 1;
        if (cfg_.autotune && l > 1) HX(tune_level(lv));
        if (paired(lv)) lv.T = 1;   // the paired launches read the T = 1 layout
        const Level0 g = level_geometry(l, lv.T, lv.block);
        lv.stride = g.stride;
        lv.ncpr = g.ncpr;
        lv.ax_ld = g.ax_ld;
        lv.vld = g.vld;
        lv.kld = round_up((lv.side - 1) * (lv.side - 1), 32);
    }
    if (budget_bytes && plan_bytes() > budget_bytes) {
        over_budget_ = true;
        return hipErrorOutOfMemory;
    }

    const ValueType vt = (ValueType)cfg_.inner_vtype;
    const size_t vs = vt_size(vt);
    for (int64_t l = 1; l <= max_level_; ++l) {
        BLevel &lv = L_[(size_t)l];
        const size_t vb = (size_t)(lv.cap * lv.vld) * inner_vector_bytes();
        for (void **p : {&lv.dinv, &lv.rhs, &lv.x, &lv.d, &lv.t}) HX(alloc(p, vb));
        HX(alloc((void **)&lv.kappa, (size_t)(lv.cap * lv.kld) * 8));
        if (stores_stencil(lv)) HX(alloc(&lv.st, (size_t)(lv.cap * 4 * lv.vld) * vs));
        if (l == 1)
            for (void **p : {&lv.g, &lv.u, &lv.en}) HX(alloc(p, vb));
        if (!stores_values(lv)) continue;   // matrix-free paired level: no index arrays or values
        HX(alloc((void **)&lv.Aj, (size_t)lv.ax_ld * 4));
        HX(alloc((void **)&lv.Arl, (size_t)lv.n * 4));
        HX(alloc(&lv.Ax, (size_t)(lv.cap * lv.ax_ld) * vs));
        // shared index arrays (values of slot 0 are placeholders until a load)
        HX(q1_assemble(vt, lv.side, lv.Aj, lv.Ax, lv.Arl, lv.stride, nullptr, lv.dinv, cfg_.omega, s_main_, lv.T));
    }
    for (BClass &c : C_) {
        for (double **p : {&c.b, &c.x, &c.d}) HX(alloc((void **)p, (size_t)(c.count * c.vld) * 8));
        if (cfg_.outer == 1) {
            HX(alloc((void **)&c.p, (size_t)(2 * c.count * c.vld) * 8));   // two direction buffers
            HX(alloc((void **)&c.q, (size_t)(c.count * c.vld) * 8));
        }
        HX(alloc((void **)&c.partials, (size_t)(c.count * 4 * c.bps) * 8));   // 64-row chunks
        if (!stores_outer_values()) continue;
        HX(alloc((void **)&c.Aj, (size_t)c.ax_ld * 4));
        HX(alloc((void **)&c.Arl, (size_t)c.n * 4));
        HX(alloc((void **)&c.Ax, (size_t)(c.count * c.ax_ld) * 8));
        HX(q1_assemble(VT_F64, c.side, c.Aj, c.Ax, c.Arl, c.stride, nullptr, nullptr, 0.0, s_main_, 1));
    }
    HX(alloc((void **)&sc_, (size_t)nslots_ * sizeof(SlotScalars)));
    HX(alloc((void **)&list_, (size_t)nslots_ * sizeof(int32_t)));
    HX(alloc((void **)&bounds_, (size_t)(kMaxClasses + 1) * sizeof(int32_t)));
    boundaries_.clear();
    for (const BClass &c : C_) boundaries_.push_back(c.first);
    dir_tiles_ = 0;
    for (const BClass &c : C_) dir_tiles_ += c.count * dir_tiles(c.side).ntiles;
    boundaries_.push_back(nslots_);
    HX(b_fill_state(sc_, nslots_, SLOT_IDLE, s_main_));
    HX(hipHostMalloc((void **)&sc_host_, (size_t)nslots_ * sizeof(SlotScalars), hipHostMallocDefault));
    host_state_.assign((size_t)nslots_, 0);
    load_ev_.assign((size_t)nslots_, nullptr);
    for (auto &e : load_ev_) HX(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HX(hipStreamSynchronize(s_main_));
    return capture();
}

BWork BatchMG::level_work(int64_t level) const {
    int32_t hi = 0;
    while (hi < (int32_t)C_.size() && C_[(size_t)hi].level >= level) ++hi;
    return BWork{list_, bounds_, 0, hi, 0};
}

BWork BatchMG::class_work(size_t k) const {
    return BWork{list_, bounds_, (int32_t)k, (int32_t)k + 1, C_[k].first};
}

BWork BatchMG::all_work() const { return BWork{list_, bounds_, 0, (int32_t)C_.size(), 0}; }

hipError_t BatchMG::enqueue_vcycle(hipStream_t s) {
    const ValueType vt = (ValueType)cfg_.inner_vtype;
    auto mat = [&](const BLevel &lv) {
        return BMat{lv.Aj, lv.Arl, lv.Ax, lv.stride, lv.ax_ld, lv.n, lv.vld, (lv.n * lv.T + lv.block - 1) / lv.block};
    };
    // levels 1..lf as one launch (one work-group per slot), the rest level by level
    const int64_t lf = cfg_.fused_levels >= 2 ? std::min<int64_t>(cfg_.fused_levels, max_level_) : 1;
    for (int64_t l = max_level_; l > lf; --l) {
        BLevel &lv = L_[(size_t)l];
        const BMat m = mat(lv);
        const BWork w = level_work(l);
        if (paired(lv) && cfg_.n_pre == 2) {
            // x = Jacobi0, d = rhs - A x: one launch
            HX(b_pair(vt, PAIR_PRE, lv, w, lv.x, lv.d, nullptr, s, pair_weights(lv)));
            HX(b_restrict(vt, lv, L_[(size_t)l - 1], w, s));
            continue;
        }
        // x = n_pre Jacobi sweeps from zero (first two fused), ending in x
        const int rest = cfg_.n_pre - 2;
        void *src = (rest % 2 == 0) ? lv.x : lv.t;
        HX(b_ell(vt, lv.T, lv.block, lv.group, ELL_JACOBI0, m, w, lv.cap, src, nullptr, lv.rhs, lv.dinv, s));
        for (int k = 0; k < rest; ++k) {
            void *dst = (src == lv.x) ? lv.t : lv.x;
            HX(b_ell(vt, lv.T, lv.block, lv.group, ELL_JACOBI, m, w, lv.cap, dst, src, lv.rhs, lv.dinv, s));
            src = dst;
        }
        HX(b_ell(vt, lv.T, lv.block, lv.group, ELL_DEFECT, m, w, lv.cap, lv.d, lv.x, lv.rhs, nullptr, s));
        HX(b_restrict(vt, lv, L_[(size_t)l - 1], w, s));
    }
    if (lf >= 2) {
        FusedCycle fc{};
        fc.top = (int)lf;
        fc.n_pre = cfg_.n_pre;
        fc.n_post = cfg_.n_post;
        for (int64_t l = 1; l <= lf; ++l) {
            const BLevel &lv = L_[(size_t)l];
            FusedLevel &f = fc.lv[l];
            f.Aj = lv.Aj;
            f.Arl = lv.Arl;
            f.Ax = lv.Ax;
            f.stride = lv.stride;
            f.ax_ld = lv.ax_ld;
            f.n = lv.n;
            f.vld = lv.vld;
            f.side = lv.side;
            f.cap = lv.cap;
            f.ncpr = lv.ncpr;
            f.T = lv.T;
            f.dinv = lv.dinv;
            f.rhs = lv.rhs;
            f.x = lv.x;
            f.d = lv.d;
            f.t = lv.t;
        }
        fc.g = L_[1].g;
        fc.u = L_[1].u;
        fc.en = L_[1].en;
        fc.konv = coarse_konv();
        fc.max_iters = (int)(4 * L_[1].n + 64);
        HX(b_coarse_vcycle(vt, fc, level_work(1), L_[1].cap, s));
    } else {
        HX(b_coarse_cg(vt, L_[1], level_work(1), s));
    }
    for (int64_t l = lf + 1; l <= max_level_; ++l) {
        BLevel &lv = L_[(size_t)l];
        const BMat m = mat(lv);
        const BWork w = level_work(l);
        const BLevel &lc = L_[(size_t)l - 1];
        if (fused_post(lv)) {
            // t = J(J(x + P e)): one launch for the correction and both sweeps
            HX(b_pair(vt, PAIR_POST, lv, w, lv.t, nullptr, lv.x, s, pair_weights(lv), &lc, level_result(lc)));
            continue;
        }
        HX(b_prolong_axpy(vt, lv, lc, level_result(lc), w, lv.x, s));
        void *src = lv.x, *dst = lv.t;
        for (int k = 0; k < cfg_.n_post; ++k) {
            HX(b_ell(vt, lv.T, lv.block, lv.group, ELL_JACOBI, m, w, lv.cap, dst, src, lv.rhs, lv.dinv, s));
            std::swap(src, dst);
        }
        if (src != lv.x) HX(b_copy(vt, lv, w, lv.x, src, s));
    }
    return hipSuccess;
}

BClassTab BatchMG::class_table() const {
    BClassTab t{};
    t.nclass = (int32_t)C_.size();
    for (size_t k = 0; k < C_.size(); ++k) {
        const BClass &c = C_[k];
        const BLevel &top = L_[(size_t)c.level];
        BClassEntry &e = t.c[k];
        e.first = c.first;
        e.n = c.n;
        e.side = c.side;
        e.vld = c.vld;
        e.stride = c.stride;
        e.ax_ld = c.ax_ld;
        e.bps = c.bps;
        e.side_magic = (uint32_t)((uint64_t(1) << 32) / (uint64_t)c.side + 1);
        e.pst = 4 * c.bps;
        e.nps = (c.n + 63) / 64;
        e.vld_top = top.vld;
        e.Aj = c.Aj;
        e.Arl = c.Arl;
        e.Ax = c.Ax;
        e.b = c.b;
        e.x = c.x;
        e.r = c.d;
        e.p = c.p;
        e.q = c.q;
        e.partials = c.partials;
        e.rhs_top = top.rhs;
        e.top64 = cfg_.inner_vtype == VT_F64 ? 1 : 0;
        e.x_top = level_result(top);
        e.kappa = top.kappa;
        e.kld = top.kld;
        e.palt = c.count * c.vld;
        const DirTiles dt = dir_tiles(c.side);
        e.tw = dt.tw;
        e.strips = dt.strips;
        e.ntiles = dt.ntiles;
    }
    return t;
}

// One outer iteration of every slot.  Defect correction (outer 0):
//   list non-idle -> r = b - A x, |r| -> check -> list active -> rhs = r/|r|
//   -> V-cycle -> x += |r| c.
// Preconditioned CG (outer 1), z = s V(r/s):
//   list new / verifying -> r = b - A x for them -> list non-idle -> check
//   (recursive |r| for running slots) -> list active -> rhs = r/|r| for the
//   slots whose r was just recomputed (s = |r|) -> V-cycle -> rho = r.z,
//   beta -> p = z + beta p with q = A p, p.q (one tiled launch) -> alpha ->
//   x += alpha p, r -= alpha q, |r|, and the next cycle's rhs = r / s' with
//   s' = |r| of this iteration's check (the new norm is not known yet). The
//   cycle is linear in its right-hand side apart from the coarse CG's
//   stopping rule, which compares absolute changes with the type's epsilon;
//   s' keeps that input within one iteration's reduction factor of unit
//   size, as r/|r| does.
hipError_t BatchMG::enqueue_iteration(hipStream_t s) {
    const int nb = (int)boundaries_.size();
    const BClassTab tab = class_table();
    const BWork w = all_work();
    int64_t blocks = 0;
    for (const BClass &c : C_) blocks += c.count * c.bps;
    const bool cg = cfg_.outer == 1, mf = cfg_.matrix_free != 0;
    if (cg) {   // CG keeps its recursive residual: only new / verifying slots need b - A x
        HX(b_compact(sc_, nslots_, boundaries_.data(), nb, list_, bounds_, LIST_UNSETTLED, s));
        HX(b_outer(OP_DEFECT_SETTLE, tab, w, blocks, sc_, s, mf));
    }
    HX(b_compact(sc_, nslots_, boundaries_.data(), nb, list_, bounds_, LIST_NONIDLE, s));
    if (!cg) HX(b_outer(OP_DEFECT_ALL, tab, w, blocks, sc_, s, mf));
    HX(b_finish(cg ? FIN_CHECK_CG : FIN_CHECK_DC, tab, w, nslots_, sc_, cfg_.tolerance, cfg_.max_outer, s));
    HX(b_compact(sc_, nslots_, boundaries_.data(), nb, list_, bounds_, LIST_ACTIVE, s));
    HX(b_outer(OP_SCALE, tab, w, blocks, sc_, s));
    HX(enqueue_vcycle(s));
    if (!cg) return b_outer(OP_ACC, tab, w, blocks, sc_, s);
    HX(b_outer(OP_DOT_RZ, tab, w, blocks, sc_, s));
    HX(b_finish(FIN_BETA, tab, w, nslots_, sc_, 0.0, 0, s));
    HX(b_dir_spmv(tab, w, dir_tiles_, sc_, s, mf));
    HX(b_finish(FIN_ALPHA, tab, w, nslots_, sc_, 0.0, 0, s));
    return b_outer(OP_UPDATE, tab, w, blocks, sc_, s);
}

hipError_t BatchMG::capture_iterations(int iters, hipGraph_t *graph, hipGraphExec_t *exec) {
    HX(hipStreamBeginCapture(s_main_, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (int i = 0; i < iters && err == hipSuccess; ++i) err = enqueue_iteration(s_main_);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(s_main_, &g);
    if (err != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return err;
    }
    HX(e2);
    *graph = g;
    return hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
}

// The full graph runs iters_per_launch iterations; a one-iteration graph
// serves launches that only a few running slots need (the tail of a batch),
// so finished work does not sit behind up to iters_per_launch - 1 empty
// iterations.
hipError_t BatchMG::capture() {
    if (!cfg_.use_graphs) return hipSuccess;
    HX(capture_iterations(cfg_.iters_per_launch, &graph_, &exec_));
    if (cfg_.iters_per_launch > 1) HX(capture_iterations(1, &graph1_, &exec1_));
    return hipSuccess;
}

hipError_t BatchMG::load(int64_t g, const double *coeff, const double *b) {
    if (!coeff || !b) return hipErrorInvalidValue;
    return load_many(1, &g, &coeff, &b, nullptr, nullptr, 0.0);
}

hipError_t BatchMG::load_seeded(int64_t g, uint64_t coeff_seed, double sigma, uint64_t rhs_seed) {
    return load_many(1, &g, nullptr, nullptr, &coeff_seed, &rhs_seed, sigma);
}

// The build runs on the load stream beside the graph launches: the slots are
// idle (their last graph work finished before the poll that retired them,
// and result() copies synchronously), and idle slots are touched by no
// kernel.  Caller data produced on the legacy default stream comes first.
hipError_t BatchMG::load_many(int64_t n, const int64_t *slots, const double *const *coeffs,
                              const double *const *bs, const uint64_t *cseeds, const uint64_t *bseeds, double sigma) {
    if (n <= 0) return hipSuccess;
    if (n > 65535 || !slots) return hipErrorInvalidValue;
    const bool seeded = coeffs == nullptr;
    if (seeded ? (!cseeds || !bseeds) : !bs) return hipErrorInvalidValue;
    std::vector<int64_t> cls((size_t)n);
    int maxl = 1;
    for (int64_t e = 0; e < n; ++e) {
        const int64_t g = slots[e];
        if (g < 0 || g >= nslots_ || host_state_[(size_t)g] != 0) return hipErrorInvalidValue;
        for (int64_t f = 0; f < e; ++f)
            if (slots[f] == g) return hipErrorInvalidValue;
        if (!seeded && (!coeffs[e] || !bs[e])) return hipErrorInvalidValue;
        size_t k = 0;
        while (!(g >= C_[k].first && g < C_[k].first + C_[k].count)) ++k;
        cls[(size_t)e] = (int64_t)k;
        maxl = std::max(maxl, (int)C_[k].level);
    }
    hipStream_t s = s_load_;
    // staging and scratch of the previous build must be free
    HX(hipEventSynchronize(build_ev_));
    if (n > build_cap_) {
        if (build_host_) (void)hipHostFree(build_host_);
        if (build_dev_) (void)hipFree(build_dev_);
        build_host_ = nullptr;
        build_dev_ = nullptr;
        build_cap_ = 0;
        HX(hipHostMalloc((void **)&build_host_, sizeof(BuildEntry) * (size_t)n, hipHostMallocDefault));
        HX(hipMalloc((void **)&build_dev_, sizeof(BuildEntry) * (size_t)n));
        build_cap_ = n;
    }
    const ValueType vt = (ValueType)cfg_.inner_vtype;
    const size_t vs = vt_size(vt);
    BuildParams P{};
    P.entries = (int32_t)n;
    P.averaging = cfg_.averaging;
    P.vt = vt;
    P.outer_values = stores_outer_values() ? 1 : 0;
    P.sigma = sigma;
    P.omega = cfg_.omega;
    for (int64_t l = 1; l <= max_level_; ++l) {
        P.side[l] = L_[(size_t)l].side;
        P.stride[l] = L_[(size_t)l].stride;
        P.T[l] = L_[(size_t)l].T;
        P.vld[l] = L_[(size_t)l].vld;
    }
    for (const BClass &c : C_) P.stride64[c.level] = c.stride;
    for (int64_t e = 0; e < n; ++e) {
        const int64_t g = slots[e];
        const BClass &c = C_[(size_t)cls[(size_t)e]];
        const int64_t ls = g - c.first;
        BuildEntry &b = build_host_[e];
        b = BuildEntry{};
        b.level = (int32_t)c.level;
        b.seeded = seeded ? 1 : 0;
        b.coeff_seed = seeded ? cseeds[e] : 0;
        b.rhs_seed = seeded ? bseeds[e] : 0;
        b.coeff_src = seeded ? nullptr : coeffs[e];
        b.b_src = seeded ? nullptr : bs[e];
        for (int64_t l = 1; l <= c.level; ++l) {
            const BLevel &lv = L_[(size_t)l];
            b.kappa[l] = lv.kappa + g * lv.kld;
            b.Ax[l] = lv.Ax ? (char *)lv.Ax + (size_t)(g * lv.ax_ld) * vs : nullptr;
            b.dinv[l] = (char *)lv.dinv + (size_t)(g * lv.vld) * inner_vector_bytes();
            b.st[l] = lv.st ? (char *)lv.st + (size_t)(g * 4 * lv.vld) * vs : nullptr;
        }
        b.Ax64 = c.Ax ? c.Ax + ls * c.ax_ld : nullptr;
        b.b = c.b + ls * c.vld;
        b.x = c.x + ls * c.vld;
    }
    if (!seeded) {
        HX(hipEventRecord(null_ev_, nullptr));
        HX(hipStreamWaitEvent(s, null_ev_, 0));
    }
    HX(hipMemcpyAsync(build_dev_, build_host_, sizeof(BuildEntry) * (size_t)n, hipMemcpyHostToDevice, s));
    HX(b_build(build_dev_, P, maxl, s));
    HX(hipEventRecord(build_ev_, s));
    for (int64_t e = 0; e < n; ++e) {
        const int64_t g = slots[e];
        HX(hipEventRecord(load_ev_[(size_t)g], s));
        host_state_[(size_t)g] = 1;
        loading_.push_back(g);
    }
    return hipSuccess;
}

hipError_t BatchMG::activate_ready(bool wait) {
    std::vector<int64_t> still, ready;
    for (int64_t g : loading_) {
        hipError_t q = wait ? hipEventSynchronize(load_ev_[(size_t)g]) : hipEventQuery(load_ev_[(size_t)g]);
        if (q == hipErrorNotReady) {
            still.push_back(g);
            continue;
        }
        HX(q);
        ready.push_back(g);
    }
    HX(b_set_states(sc_, ready.data(), (int64_t)ready.size(), SLOT_NEW, s_main_));
    for (int64_t g : ready) host_state_[(size_t)g] = 2;
    running_ += (int64_t)ready.size();
    loading_.swap(still);
    return hipSuccess;
}

hipError_t BatchMG::step() {
    HX(activate_ready(false));
    // the first launch of a batch (none since the pool last ran empty),
    // sized from the last batch
    const bool first = cfg_.first_launch_from_history && !batch_launched_ && running_ > 0 && first_iters_ > 0;
    if (running_ > 0) batch_launched_ = true;
    // a tail: few slots still running and none waiting to join
    const bool tail = cfg_.iters_per_launch > 1 && loading_.empty() && running_ <= std::max<int64_t>(1, nslots_ / 32);
    if (cfg_.use_graphs) {
        HX(hipGraphLaunch(first ? execf_ : (tail ? exec1_ : exec_), s_main_));
    } else {
        const int n = first ? first_iters_ : (tail ? 1 : cfg_.iters_per_launch);
        for (int i = 0; i < n; ++i) HX(enqueue_iteration(s_main_));
    }
    ++launches_;
    return hipSuccess;
}

// Once a batch has retired entirely: the iteration count its first launch
// would have needed, and (with graphs) that launch captured now, outside the
// next batch's time.
hipError_t BatchMG::note_batch_end() {
    const int32_t k = std::min<int32_t>(batch_max_iters_ + 2, cfg_.iters_per_launch);
    batch_max_iters_ = 0;
    if (k == first_iters_) return hipSuccess;
    if (execf_) (void)hipGraphExecDestroy(execf_);
    if (graphf_) (void)hipGraphDestroy(graphf_);
    execf_ = nullptr;
    graphf_ = nullptr;
    first_iters_ = 0;
    if (cfg_.use_graphs) HX(capture_iterations(k, &graphf_, &execf_));
    first_iters_ = k;
    return hipSuccess;
}

hipError_t BatchMG::poll(std::vector<SlotScalars> &out, std::vector<int64_t> *retired) {
    HX(hipMemcpyAsync(sc_host_, sc_, (size_t)nslots_ * sizeof(SlotScalars), hipMemcpyDeviceToHost, s_main_));
    HX(hipStreamSynchronize(s_main_));
    out.assign(sc_host_, sc_host_ + nslots_);
    if (retired) retired->clear();
    bool any = false;
    for (int64_t g = 0; g < nslots_; ++g) {
        if (host_state_[(size_t)g] == 2 && sc_host_[g].state == SLOT_IDLE) {
            host_state_[(size_t)g] = 0;
            --running_;
            batch_max_iters_ = std::max(batch_max_iters_, sc_host_[g].iters);
            any = true;
            if (retired) retired->push_back(g);
        }
    }
    if (running_ == 0 && loading_.empty()) {
        batch_launched_ = false;
        if (cfg_.first_launch_from_history && any) HX(note_batch_end());
    }
    return hipSuccess;
}

const double *BatchMG::solution(int64_t g, int64_t *n) const {
    for (const BClass &c : C_)
        if (g >= c.first && g < c.first + c.count) {
            if (n) *n = c.n;
            return c.x + (g - c.first) * c.vld;
        }
    return nullptr;
}

hipError_t BatchMG::result(int64_t g, double *dst) {
    if (g < 0 || g >= nslots_ || !dst) return hipErrorInvalidValue;
    for (const BClass &c : C_)
        if (g >= c.first && g < c.first + c.count) {
            HX(hipMemcpyAsync(dst, c.x + (g - c.first) * c.vld, (size_t)c.n * 8, hipMemcpyDeviceToDevice, s_main_));
            return hipStreamSynchronize(s_main_);
        }
    return hipErrorInvalidValue;
}

hipError_t BatchMG::time_level(int64_t level, int reps, double us[5]) {
    if (level < 2 || level > max_level_ || reps < 1 || !us) return hipErrorInvalidValue;
    HX(hipStreamSynchronize(s_load_));
    HX(hipStreamSynchronize(s_main_));
    for (int k = 0; k < 5; ++k) us[k] = 0.0;
    // every slot listed; bounds[k] = slots below class boundary k
    std::vector<int32_t> hl((size_t)nslots_), hb((size_t)kMaxClasses + 1, 0);
    for (int64_t g = 0; g < nslots_; ++g) hl[(size_t)g] = (int32_t)g;
    for (size_t k = 0; k < boundaries_.size(); ++k) hb[k] = (int32_t)boundaries_[k];
    int32_t *dl = nullptr, *db = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc((void **)&dl, hl.size() * sizeof(int32_t));
    if (err == hipSuccess) err = hipMalloc((void **)&db, hb.size() * sizeof(int32_t));
    if (err == hipSuccess) err = hipMemcpy(dl, hl.data(), hl.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(db, hb.data(), hb.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    BLevel &lv = L_[(size_t)level], &lc = L_[(size_t)level - 1];
    int32_t hi = 0;
    while (hi < (int32_t)C_.size() && C_[(size_t)hi].level >= level) ++hi;
    const BWork w{dl, db, 0, hi, 0}, wa{dl, db, 0, (int32_t)C_.size(), 0};
    const ValueType vt = (ValueType)cfg_.inner_vtype;
    auto timeit = [&](auto launch, double *out) {
        if (err != hipSuccess) return;
        err = launch();
        if (err == hipSuccess) err = hipEventRecord(e0, s_main_);
        for (int r = 0; r < reps && err == hipSuccess; ++r) err = launch();
        if (err == hipSuccess) err = hipEventRecord(e1, s_main_);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err == hipSuccess) *out = 1000.0 * ms / reps;
    };
    const bool mf = cfg_.matrix_free != 0;
    const int wm = pair_weights(lv);
    if (paired(lv)) timeit([&]() { return b_pair(vt, PAIR_PRE, lv, w, lv.x, lv.d, nullptr, s_main_, wm); }, &us[0]);
    timeit([&]() { return b_restrict(vt, lv, lc, w, s_main_); }, &us[1]);
    if (fused_post(lv))
        timeit([&]() { return b_pair(vt, PAIR_POST, lv, w, lv.t, nullptr, lv.x, s_main_, wm, &lc, lc.x); }, &us[3]);
    else
        timeit([&]() { return b_prolong_axpy(vt, lv, lc, lc.x, w, lv.t, s_main_); }, &us[2]);
    if (paired(lv) && !fused_post(lv))
        timeit([&]() { return b_pair(vt, PAIR_POST, lv, w, lv.x, nullptr, lv.t, s_main_, wm); }, &us[3]);
    if (cfg_.outer == 1) {
        const BClassTab tab = class_table();
        timeit([&]() { return b_dir_spmv(tab, wa, dir_tiles_, sc_, s_main_, mf); }, &us[4]);
    }
    (void)hipStreamSynchronize(s_main_);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (dl) (void)hipFree(dl);
    if (db) (void)hipFree(db);
    return err;
}

hipError_t BatchMG::run_to_completion(int64_t *launches) {
    HX(activate_ready(true));
    int64_t k = 0;
    std::vector<SlotScalars> st;
    while (running_ > 0) {
        HX(step());
        ++k;
        HX(poll(st));
    }
    if (launches) *launches = k;
    return hipSuccess;
}

}  // namespace hipx
}  // namespace honei
