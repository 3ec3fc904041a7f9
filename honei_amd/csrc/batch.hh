// batch.hh -- many independent Q1 systems of uneven sizes solved together by
// mixed-precision multigrid defect correction (FP32 or BF16-stored V-cycle
// inside, FP64 defect / update / norm outside), one hipGraph launch per
// group of outer iterations.
//
// Slots.  The pool is a list of classes; class c holds count_c slots whose
// finest grid is level L_c (side 2^L_c + 1).  Classes are kept in descending
// level order, so global slot g = first_c + j is present on every level
// l <= L_c, and the slots present on level l are exactly the prefix
// [0, cap_l) with cap_l = sum of count_c over classes with L_c >= l.  Each
// level therefore stores its batch as [cap_l][...] arrays indexed by the
// global slot; index arrays (Aj, Arl) are shared by all slots of a level, and
// values (Ax, dinv) and vectors are per slot.  A V-cycle of the batch runs the
// same launches as a single-system V-cycle, each over the whole prefix: a
// slot whose finest level is l enters the cycle at l with its scaled outer
// defect as right-hand side; deeper slots enter with the restricted defect.
//
// Per-slot state lives on the device (SlotScalars): every kernel returns at
// once for idle slots, and a check kernel at the start of each outer
// iteration forms each slot's FP64 defect norm, decides convergence and
// retires the slot.  Work admitted while others run is assembled on a
// separate stream into an idle slot and activated between graph launches.
//
// Numerics per slot follow MGHierarchy::solve_mixed with n_max_iter = 1,
// initial_zero, zero_inner_start and no inner convergence check (the lean
// inner cycle), so with threads-per-row 1 a slot's result is bit-identical
// to that single-system solve of the same system.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "solvers.hh"

namespace honei {
namespace hipx {

// VERIFY: the recursive CG residual met the tolerance; the next iteration
// recomputes b - A x before the slot retires.
enum SlotState : int32_t { SLOT_IDLE = 0, SLOT_ACTIVE = 1, SLOT_NEW = 2, SLOT_VERIFY = 3 };
enum SlotStatus : int32_t { SLOT_OK = 0, SLOT_PENDING = -1, SLOT_NOT_CONVERGED = 5 };

struct SlotScalars {
    double initial;   // FP64 defect norm at the start vector
    double norm;      // latest FP64 defect norm
    double scale;     // inner right-hand side scaling of the current cycle
    double inv;       // 1 / scale
    double rel;       // norm / initial
    double rho;       // CG: r.z of the previous iteration (0: restart)
    double beta;      // CG: direction update factor of this iteration
    double alpha;     // CG: step length of this iteration
    // CG: the scaling OP_UPDATE gives the next cycle's right-hand side
    // (r_new * inv_w, z then taken as scale_w * V): 1/|r| and |r| of the latest check
    double scale_w, inv_w;
    int32_t state;    // SlotState
    int32_t iters;    // outer iterations (V-cycles) applied
    int32_t status;   // SlotStatus
    int32_t pcur;     // CG: which of the slot's two direction buffers holds p (low bit)
    int32_t fresh;    // 1: r was (re)computed from b - A x at this check, OP_SCALE writes the rhs
    int32_t pad;
};

constexpr int kMaxClasses = 16;

// Paired row passes (b_pair): tiles of at most kPairWidth columns (62 + a
// halo column each side fill a wavefront) and kPairRows rows (kPairTileRows
// with the halo, one wavefront per row: 1024-thread work-groups), on levels
// from kPairMinSide^2 nodes up; from kPairTallMinSide^2 nodes up every thread
// takes two grid rows (tiles of 2 kPairTileRows - 2 rows).  The CG direction
// launch tiles the finest grids in the same column strips and bands of
// kDirRows rows (kDirNodes grid rows per thread).
enum PairKind { PAIR_PRE = 0, PAIR_POST = 1 };
constexpr int64_t kPairWidth = 62;
constexpr int kPairTileRows = 16, kPairRows = kPairTileRows - 2;
constexpr int64_t kPairMinSide = 33, kPairTallMinSide = 129;
constexpr int kDirNodes = 2, kDirRows = kPairTileRows * kDirNodes - 2;

struct BatchConfig {
    // VT_F32 or VT_BF16 matrix values with FP32 vectors (mixed precision),
    // or VT_F64 values and vectors (the same solver entirely in FP64)
    int32_t inner_vtype = VT_F32;
    int32_t n_pre = 2, n_post = 2;   // Jacobi sweeps (n_pre >= 2)
    double omega = 0.8;              // Jacobi damping
    double tolerance = 1e-8;         // relative FP64 defect
    int32_t max_outer = 200;
    int32_t averaging = 1;           // coefficient coarsening: 0 arith, 1 geometric, 2 harmonic
    int32_t iters_per_launch = 2;    // outer iterations per graph launch
    int32_t autotune = 1;            // 0: threads-per-row 1, block 256 everywhere
    // Outer iteration: 0 defect correction x += V(b - A x) (the reference's
    // MIXED driver), 1 conjugate gradients in FP64 preconditioned by the
    // V-cycle (robust where the rediscretised hierarchy over-corrects).
    int32_t outer = 1;
    // Levels 1..fused_levels of every V-cycle in one launch (one work-group
    // per slot); 0 or 1: level by level.
    int32_t fused_levels = 4;
    // 1: each step is one graph launch; 0: the same kernels enqueued on the
    // stream at every step (for tools that cannot follow graph launches).
    int32_t use_graphs = 1;
    // 1: on levels of side >= kPairMinSide (T = 1 there), the Jacobi0 +
    // defect and the two post-smoothing sweeps each run as one launch that
    // reads the matrix once (n_pre == 2 / n_post == 2); 0: one launch per pass.
    int32_t pair_sweeps = 1;
    // Where the paired launches and the FP64 outer row sums take each row's
    // nine weights from (all forms give the stored matrix's bits):
    //   0: the stored values (9 per row);
    //   1: formed from the slot's cell coefficients (8 bytes per node) with
    //      the assembly's own expressions, paired launches and FP64 outer;
    //   2: paired launches gather them from the distinct weights stored
    //      four per node, interleaved (q1_stencil_node; one per cell, edge
    //      and node), built with the hierarchy; FP64 outer as 1.
    int32_t matrix_free = 1;
    // 1: a launch that finds no slot running (the first of a batch) runs as
    // many outer iterations as the last such batch needed until its last
    // slot retired (its largest iteration count + 2: the check that meets
    // the tolerance and the verification), at most iters_per_launch; later
    // launches unchanged.  0: every launch runs iters_per_launch.
    int32_t first_launch_from_history = 0;
};

// Inner level l of the pool: slots [0, cap) present.
struct BLevel {
    int64_t level = 0, side = 0, n = 0, cap = 0, vld = 0;
    int T = 1, block = 256, group = 1;   // lanes per row, work-group size, slots per thread
    int64_t stride = 0, ncpr = 0, ax_ld = 0;
    int32_t *Aj = nullptr;
    uint32_t *Arl = nullptr;
    void *Ax = nullptr;
    // vectors of the inner type (float, or double for an FP64 inner solve)
    void *dinv = nullptr, *rhs = nullptr, *x = nullptr, *d = nullptr, *t = nullptr;
    void *g = nullptr, *u = nullptr, *en = nullptr;   // coarse CG scratch (level 1 only)
    double *kappa = nullptr;   // cell coefficients [cap][kld] ((side-1)^2 per slot)
    int64_t kld = 0;
    void *st = nullptr;        // distinct weights [cap][vld][4] of the value type (matrix_free 2)
};

// Class c: slots [first, first + count), finest level `level`, FP64 outer data.
struct BClass {
    int64_t level = 0, side = 0, n = 0, first = 0, count = 0, vld = 0;
    int64_t stride = 0, ax_ld = 0, bps = 0;
    int32_t *Aj = nullptr;
    uint32_t *Arl = nullptr;
    double *Ax = nullptr, *b = nullptr, *x = nullptr, *d = nullptr, *partials = nullptr;
    double *p = nullptr, *q = nullptr;   // CG direction (two buffers, [2][count][vld]) and A p
};

// Threads-per-row / block-size choice for one level (autotuned, cached per
// level and value type for the process).
struct LaunchShape {
    int T = 1, block = 256, group = 1;
    double us = 0.0;       // measured time of one batched Jacobi sweep
    int64_t slots = 0;     // slots of the tuning batch
    // paired launches (b_pair) on the same batch, T = 1, with the weights
    // from the stored values [0], the cell coefficients [1], the arrays of
    // distinct weights [2]
    double pre_us[3] = {0.0, 0.0, 0.0}, post_us[3] = {0.0, 0.0, 0.0};
};

// Batched kernels (batch_kernels.hip).
//
// A launch covers the active slots list[a], a in [bounds[lo], bounds[hi]):
// `list` holds the non-idle slots in ascending order and bounds[k] counts the
// active slots below class boundary k (boundary k = first slot of class k,
// boundary nclass = nslots), both written on the device by b_compact.  Every
// batched kernel runs on a fixed grid and walks the active slots' row blocks
// in contiguous chunks, so idle slots cost nothing and the launch
// configuration never changes (graph-capturable).
struct BWork {
    const int32_t *list;
    const int32_t *bounds;
    int32_t lo, hi;
    int64_t base;   // slot of local index 0 in the launch's arrays
};

// Per-class data of the FP64 outer launches, which cover the active slots of
// every class in one launch (logical block = 256 rows of one slot).
struct BClassEntry {
    int64_t first, n, side, vld, stride, ax_ld, bps, vld_top;
    uint32_t side_magic;   // floor(2^32 / side) + 1: v / side = umulhi(v, side_magic) for v side < 2^32
    int64_t pst;   // partial stride per slot: 4 bps
    int64_t nps;   // partials per slot of the CG reductions: one per 64-row chunk
    const int32_t *Aj;
    const uint32_t *Arl;
    const double *Ax, *b;
    double *x, *r, *p, *q, *partials;
    void *rhs_top;      // inner right-hand side of the class's finest level
    const void *x_top;  // inner solution of the class's finest level
    int32_t top64;      // 1: the inner vectors are double, 0: float
    const double *kappa;   // cell coefficients of the finest level, [global slot][kld]
    int64_t kld;
    int64_t palt;          // CG: offset of the second direction buffer from p (count * vld)
    int32_t tw, strips;    // tiles of k_b_dir_spmv: e.tw columns, `strips` tiles across
    int64_t ntiles;        // tiles per slot (strips * row bands of kDirRows)
};
struct BClassTab {
    int32_t nclass;
    BClassEntry c[kMaxClasses];
};

enum OuterOp : int32_t {
    OP_DEFECT_ALL = 0,     // r = b - A x (+ |r|^2) for every listed slot
    OP_DEFECT_SETTLE = 1,  // same op; CG lists only new / verifying slots for it
    OP_SCALE = 2,          // rhs_top = r / |r| in the inner precision (slots with `fresh`; CG's
                           // running slots get theirs from OP_UPDATE)
    OP_ACC = 3,            // x = (first cycle ? 0 : x) + |r| * x_top   (defect correction)
    OP_DOT_RZ = 4,         // partial r.z, z = |r| * x_top
    // (p = z + beta p with q = A p, p.q: b_dir_spmv)
    OP_UPDATE = 7          // x += alpha p, r -= alpha q, partial |r|^2, rhs_top = r * inv_w
};
enum FinishOp : int32_t { FIN_CHECK_DC = 0, FIN_CHECK_CG = 1, FIN_BETA = 2, FIN_ALPHA = 3 };

struct BBoundaries {
    int32_t n;
    int64_t v[kMaxClasses + 1];
};

struct BMat {
    const int32_t *Aj;
    const uint32_t *Arl;
    const void *Ax;
    int64_t stride, ax_ld, n, vld, bps;
    const double *kappa;   // cell coefficients [slot][kld] (matrix-free launches)
    int64_t kld;
    const void *st;        // distinct weights [slot][vld][4] (stencil-array launches)
};

// ---- bulk hierarchy build (many slots per launch) ----------------------------
constexpr int kMaxLevel = 13;

// One slot to build: coefficient and right-hand side from device arrays or
// drawn from seeds; every level's per-slot values, the FP64 outer operator,
// b and the start vector are written.
struct BuildEntry {
    int32_t level;                // finest level of the slot
    int32_t seeded;               // 1: coefficient exp(sigma N(0,1)), b ~ U[-1,1) from the seeds
    uint64_t coeff_seed, rhs_seed;
    const double *coeff_src, *b_src;
    double *kappa[kMaxLevel + 1]; // coefficient of level l ((2^l)^2 cells) of this slot; kappa[level] finest
    void *Ax[kMaxLevel + 1];      // this slot's inner values per level
    void *dinv[kMaxLevel + 1];    // this slot's omega / diag per level (inner vector type)
    void *st[kMaxLevel + 1];      // this slot's distinct-weight arrays per level, or null
    double *Ax64, *b, *x;         // FP64 outer values, right-hand side, start vector
};
struct BuildParams {
    int32_t entries, averaging, vt;
    int32_t outer_values;         // 0: the pool stores no FP64 outer values (matrix-free outer)
    double sigma, omega;
    int64_t side[kMaxLevel + 1], stride[kMaxLevel + 1], stride64[kMaxLevel + 1], vld[kMaxLevel + 1];
    int32_t T[kMaxLevel + 1];
};

class BatchMG {
  public:
    BatchMG() = default;
    ~BatchMG();
    BatchMG(const BatchMG &) = delete;
    BatchMG &operator=(const BatchMG &) = delete;

    // levels[c] (2..13) with counts[c] >= 1; any order (sorted internally,
    // class_of maps the caller's class index to the internal one).
    hipError_t create(const int64_t *levels, const int64_t *counts, int nclass, const BatchConfig &cfg,
                      uint64_t budget_bytes);
    uint64_t bytes() const { return bytes_; }
    bool over_budget() const { return over_budget_; }
    // Device bytes the pool needs (after launch shapes are chosen).
    uint64_t plan_bytes() const;
    int64_t nslots() const { return nslots_; }
    int64_t slot(int cls, int64_t j) const;   // -1 if out of range
    int64_t slot_level(int64_t g) const;

    // Assemble the hierarchy of slot g from a per-cell coefficient
    // ((side-1)^2 doubles, device) and copy the right-hand side (side^2
    // doubles, device); start vector 0 with boundary rows set to b.  Runs on
    // the load stream; the slot joins the next launch after the work is done.
    hipError_t load(int64_t g, const double *coeff_dev, const double *b_dev);
    // Same with the coefficient exp(sigma N(0,1)) and b uniform in [-1, 1)
    // drawn on the device from the seeds (synthetic requests).
    hipError_t load_seeded(int64_t g, uint64_t coeff_seed, double sigma, uint64_t rhs_seed);
    // Many slots in one bulk build (a few launches for all of them): from
    // device arrays (coeffs/bs non-null) or from seeds (cseeds/bseeds).
    hipError_t load_many(int64_t n, const int64_t *slots, const double *const *coeffs, const double *const *bs,
                         const uint64_t *cseeds, const uint64_t *bseeds, double sigma);
    // Activate finished loads, then one graph launch (iters_per_launch outer
    // iterations of every active slot).  Non-blocking.
    hipError_t step();
    // Wait for the launched work and copy every slot's scalars to host;
    // `retired` (optional) receives the slots that finished since the last poll.
    hipError_t poll(std::vector<SlotScalars> &out, std::vector<int64_t> *retired = nullptr);
    // x of slot g (FP64, side^2) into dst (device); synchronous.
    hipError_t result(int64_t g, double *dst);
    // Queue finished loads for the next launch (wait: block until every load
    // issued so far is done).
    hipError_t activate_ready(bool wait);
    // Load-free driver: activate everything loaded and step until no slot is
    // active.  launches: graph launches issued.
    hipError_t run_to_completion(int64_t *launches);
    // Slot is loading or running (host view, updated by poll).
    bool busy(int64_t g) const { return host_state_[(size_t)g] != 0; }
    bool loading(int64_t g) const { return host_state_[(size_t)g] == 1; }
    int64_t running() const { return running_; }
    // stream of the bulk builds (work enqueued on it before load_many runs first)
    hipStream_t load_stream() const { return s_load_; }
    // FP64 solution of slot g (device, side^2 values; *n receives side^2), or null
    const double *solution(int64_t g, int64_t *n) const;
    int64_t loading_count() const { return (int64_t)loading_.size(); }
    int64_t launches() const { return launches_; }
    const std::vector<BLevel> &levels() const { return L_; }
    const BatchConfig &config() const { return cfg_; }

    // Time one level's V-cycle launches over every slot present on the level
    // (the step's launch shape at full activity), HIP events on the pool's
    // stream, reps launches after one warm-up: us[0] paired pre-smoothing +
    // defect, us[1] restriction to level - 1, us[2] prolongation + correction
    // from level - 1, us[3] paired post-smoothing (0 where the level does not
    // pair; with the prolongation fused in, us[2] is 0 and us[3] covers
    // both), us[4] the FP64 p = z + beta p with q = A p, p.q (b_dir_spmv)
    // over every slot of every class.
    // Overwrites the pool's vectors: call between steps only (a load
    // rebuilds them, and results must be copied out first).
    hipError_t time_level(int64_t level, int reps, double us[5]);

    // level runs the paired launches (b_pair) of its smoothing passes
    bool paired(const BLevel &lv) const {
        return cfg_.pair_sweeps && lv.side >= kPairMinSide && (cfg_.n_pre == 2 || cfg_.n_post == 2);
    }
    // Some kernel reads level l's stored values: every level except those
    // whose smoothing passes all run as matrix-free paired launches outside
    // the fused coarse launch.
    bool stores_values(const BLevel &lv) const {
        const int64_t lf = cfg_.fused_levels >= 2 ? std::min<int64_t>(cfg_.fused_levels, max_level_) : 1;
        const bool all_paired = paired(lv) && cfg_.n_pre == 2 && (cfg_.n_post == 2 || cfg_.n_post == 0);
        return !(cfg_.matrix_free && all_paired && lv.level > lf);
    }
    // the up-leg of the level is one paired post-smoothing launch with the
    // prolongation from the level below fused in (b_pair with xc); it leaves
    // the level's result in t, every other level's in x
    bool fused_post(const BLevel &lv) const {
        const int64_t lf = cfg_.fused_levels >= 2 ? std::min<int64_t>(cfg_.fused_levels, max_level_) : 1;
        return paired(lv) && cfg_.n_post == 2 && lv.level > lf;
    }
    void *level_result(const BLevel &lv) const { return fused_post(lv) ? lv.t : lv.x; }
    // the level's paired launches gather the weights from the stencil arrays
    bool stores_stencil(const BLevel &lv) const { return cfg_.matrix_free == 2 && !stores_values(lv); }
    // weight source of the level's paired launches (b_pair's wmode)
    int pair_weights(const BLevel &lv) const {
        return cfg_.matrix_free == 0 ? 0 : (stores_stencil(lv) ? 2 : 1);
    }
    // the FP64 outer values are read only when the outer row sums do not
    // form them from the cell coefficients
    bool stores_outer_values() const { return !cfg_.matrix_free; }
    // bytes of one inner vector element (double for an FP64 inner solve)
    size_t inner_vector_bytes() const { return cfg_.inner_vtype == VT_F64 ? 8 : 4; }
    // konv_rad of the coarsest-level CG: the inner type's epsilon
    double coarse_konv() const {
        return cfg_.inner_vtype == VT_F64 ? 2.220446049250313e-16 : 1.1920928955078125e-07;
    }
  private:
    BatchConfig cfg_;
    std::vector<BLevel> L_;          // indexed by level (0 unused)
    std::vector<BClass> C_;          // descending level
    std::vector<int> class_of_;      // caller class -> internal class
    int64_t nslots_ = 0, max_level_ = 0;
    uint64_t bytes_ = 0;
    std::vector<void *> allocs_;
    SlotScalars *sc_ = nullptr;
    SlotScalars *sc_host_ = nullptr;   // pinned
    int32_t *list_ = nullptr;          // active slots (device, b_compact)
    int32_t *bounds_ = nullptr;        // active counts below each class boundary
    std::vector<int64_t> boundaries_;  // class firsts, then nslots

    hipStream_t s_main_ = nullptr, s_load_ = nullptr;
    hipGraph_t graph_ = nullptr, graph1_ = nullptr;      // iters_per_launch / one iteration
    int64_t dir_tiles_ = 0;   // k_b_dir_spmv tiles of all slots of all classes
    hipGraphExec_t exec_ = nullptr, exec1_ = nullptr;
    // first_launch_from_history: the first launch's iteration count (0: none
    // known yet), its graph, the largest iteration count of the slots retired
    // since the pool last ran empty, and whether it launched since then
    int32_t first_iters_ = 0, batch_max_iters_ = 0;
    bool batch_launched_ = false;
    hipGraph_t graphf_ = nullptr;
    hipGraphExec_t execf_ = nullptr;
    std::vector<int> host_state_;      // 0 free, 1 loading, 2 running
    std::vector<int64_t> loading_;
    int64_t running_ = 0, launches_ = 0;
    bool over_budget_ = false;
    std::vector<hipEvent_t> load_ev_;
    hipEvent_t null_ev_ = nullptr;

    hipError_t alloc(void **p, size_t bytes);
    // bulk build state: device entry table, pinned staging
    BuildEntry *build_dev_ = nullptr, *build_host_ = nullptr;
    int64_t build_cap_ = 0;
    hipEvent_t build_ev_ = nullptr;
    hipError_t tune_level(BLevel &lv);
    hipError_t capture();
    hipError_t capture_iterations(int iters, hipGraph_t *graph, hipGraphExec_t *exec);
    hipError_t note_batch_end();
    hipError_t enqueue_iteration(hipStream_t s);
    hipError_t enqueue_vcycle(hipStream_t s);
    BWork level_work(int64_t level) const;
    BWork class_work(size_t k) const;
    BWork all_work() const;
    BClassTab class_table() const;
    void release();
};

// Shape chosen for (device, level, value type), if the autotuner ran.
bool batch_tuned_shape(int dev, int64_t level, int vt, LaunchShape *out);

// grid of `block`-thread work-groups for up to `items` logical blocks
int batch_grid(int64_t items, int block);

hipError_t b_build(const BuildEntry *entries_dev, const BuildParams &P, int max_level, hipStream_t s);

// Which slots a compaction lists.
enum ListMode : int { LIST_NONIDLE = 0, LIST_ACTIVE = 1, LIST_UNSETTLED = 2 /* NEW or VERIFY */ };
hipError_t b_compact(const SlotScalars *sc, int64_t nslots, const int64_t *boundaries, int nb, int32_t *list,
                     int32_t *bounds, int mode, hipStream_t s);
hipError_t b_ell(ValueType vt, int T, int block, int group, EllMode mode, const BMat &m, const BWork &w,
                 int64_t max_slots, void *out, const void *x, const void *rhs, const void *dinv, hipStream_t s);
// FP64 outer work over the active slots of all classes (w: lo = 0, hi = nclass).
// max_blocks: 256-row blocks of all slots of all classes.
// mf: row sums from the cell coefficients (BClassEntry::kappa) instead of
// the stored FP64 values.
hipError_t b_outer(OuterOp op, const BClassTab &tab, const BWork &w, int64_t max_blocks, const SlotScalars *sc,
                   hipStream_t s, bool mf = false);
// CG: p = z + beta p into the slot's other direction buffer and q = A p with
// one partial p.q per tile (k_b_dir_spmv), over the active slots of all
// classes; max_tiles: the tiles of all slots of all classes.
hipError_t b_dir_spmv(const BClassTab &tab, const BWork &w, int64_t max_tiles, const SlotScalars *sc, hipStream_t s,
                      bool mf);
// One block per listed slot: reduce its partials and update its scalars.
hipError_t b_finish(FinishOp op, const BClassTab &tab, const BWork &w, int64_t max_slots, SlotScalars *sc,
                    double tol, int32_t max_outer, hipStream_t s);
hipError_t b_restrict(ValueType vt, const BLevel &fine, const BLevel &coarse, const BWork &w, hipStream_t s);
// out = fine.x + P xc (out: fine.x or fine.t; xc: the coarse level's
// result, coarse.x or coarse.t)
hipError_t b_prolong_axpy(ValueType vt, const BLevel &fine, const BLevel &coarse, const void *xc, const BWork &w,
                          void *out, hipStream_t s);
// Two dependent row passes of a T = 1 Q1 level in one launch (k_b_pair):
// PAIR_PRE writes x = Jacobi0 to out0 and d = rhs - A x to out1; PAIR_POST
// writes two Jacobi sweeps of xin to out0 (xin != out0).  Bit-identical to
// the separate b_ell launches.
// wmode: the weights from the stored values (0), from lv.kappa (1), from the
// distinct-weight arrays lv.st (2): the same bits in every form.
// xc (PAIR_POST, with `coarse`): the prolongation fused in, out0 =
// J(J(xin + P xc)), the same bits as b_prolong_axpy into a third vector
// followed by the unfused pair (xc: the coarse level's result).
hipError_t b_pair(ValueType vt, int pair, const BLevel &lv, const BWork &w, void *out0, void *out1,
                  const void *xin, hipStream_t s, int wmode = 0, const BLevel *coarse = nullptr,
                  const void *xc = nullptr);
// The distinct-weight arrays of `count` slots of a side x side level from
// their cell coefficients (kappa + g kld, st + g 4 vld).
hipError_t b_stencil_fill(ValueType vt, int64_t side, const double *kappa, int64_t kld, void *st, int64_t vld,
                          int64_t count, hipStream_t s);
hipError_t b_copy(ValueType vt, const BLevel &lv, const BWork &w, void *dst, const void *src, hipStream_t s);
hipError_t b_coarse_cg(ValueType vt, const BLevel &lv, const BWork &w, hipStream_t s);
// Slot ids of one state-setting launch (kernel argument).
constexpr int kMaxSlotIds = 256;
struct SlotIds {
    int32_t n;
    int32_t id[kMaxSlotIds];
};
// The coarse end of the V-cycle in one launch: levels 1..top of every
// listed slot by one work-group per slot (top = the slot's finest level if
// that is below `top`), the same sweeps, transfers and coarse CG as the
// level-by-level launches, in the same order and arithmetic.
constexpr int kMaxFusedLevels = 6;
struct FusedLevel {
    const int32_t *Aj;
    const uint32_t *Arl;
    const void *Ax;
    int64_t stride, ax_ld, n, vld, side, cap, ncpr;
    int T;
    void *dinv, *rhs, *x, *d, *t;   // inner vector type
};
// Byte offsets of a level's arrays in the LDS form of the fused cycle.
struct FusedLds {
    int32_t aj, arl, ax, dinv, rhs, x, d, t;
};
struct FusedCycle {
    int top, n_pre, n_post;
    FusedLevel lv[kMaxFusedLevels + 1];   // index = level (1..top)
    void *g, *u, *en;                     // level-1 CG scratch
    double konv;
    int max_iters;
    // LDS form (set by b_coarse_vcycle): per level, and the CG scratch
    FusedLds lo[kMaxFusedLevels + 1];
    int32_t og, ou, oen;
};
hipError_t b_coarse_vcycle(ValueType vt, const FusedCycle &fc, const BWork &w, int64_t max_slots, hipStream_t s);

// Fresh scalars in `state` for slots[0..n), one launch per 256 ids.
hipError_t b_set_states(SlotScalars *sc, const int64_t *slots, int64_t n, int32_t state, hipStream_t s);
hipError_t b_fill_state(SlotScalars *sc, int64_t nslots, int32_t state, hipStream_t s);

}  // namespace hipx
}  // namespace honei
