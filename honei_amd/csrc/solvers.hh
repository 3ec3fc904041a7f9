// solvers.hh -- device solver building blocks and the host-side solver
// drivers (CG, multigrid FIXED / MIXED) of the MI355X backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "backend.hh"

// Return the first failing HIP status of a host-side driver step.
#define HX(call)                                  \
    do {                                          \
        hipError_t _e = (call);                   \
        if (_e != hipSuccess) return _e;          \
    } while (0)

namespace honei {
namespace hipx {

// Device-resident CG state (math/cg.hh:63-133 scalars).
struct CGScalars {
    double alpha, lambda, talpha, initial, current, eps_rel;
    // the scalar an iteration carries to the next (CG: |r|^2, PCG: r.z) of
    // iteration k in carry[k & 1] (read by the three-kernel iteration while
    // block 0 writes the next one)
    double carry[2];
    int done, iters, max_iters, converged;   // converged: set by the PCG-Jacobi test only
};

hipError_t cg_init(ValueType vt, CGScalars *sc, double eps_rel, int max_iters, hipStream_t s);

// Scalars of the multigrid-preconditioned CG of MGHierarchy::solve_pcg, kept
// on the device: the host enqueues iterations in chunks and reads this
// record once per chunk.  Once `done` is set no kernel of an iteration
// changes x, r, p or a scalar any more.
struct MgPcgScalars {
    double r2;        // ||r||^2 of the recursive residual
    double rz;        // r.z
    double pq;        // p.q
    double r2_true;   // ||b - A x||^2 of a verification
    double initial, nrm, rho_prev, beta, alpha, malpha, tol;
    int32_t done;     // converged, broken down, non-finite or out of iterations
    int32_t verify;   // the recursive residual met the tolerance: the true defect decides
    int32_t iters, cap;
};
// Steps of one iteration (see MGHierarchy::solve_pcg): init after the first
// defect; scale (rhs = r / |r|, FP32 or FP64 inner right-hand side); z =
// |r| x (FP32 or FP64 inner solution); beta (counts the iteration); p update;
// alpha; check (after the update); confirm (after the gated true defect).
enum MgPcgStep : int { MGPCG_INIT = 0, MGPCG_BETA = 1, MGPCG_ALPHA = 2, MGPCG_CHECK = 3, MGPCG_CONFIRM = 4 };
hipError_t mgpcg_scalar(int step, MgPcgScalars *sc, double tol, int cap, hipStream_t s);
hipError_t mgpcg_scale(ValueType inner, void *rhs, const double *r, int64_t n, const MgPcgScalars *sc,
                       hipStream_t s);
hipError_t mgpcg_z(ValueType inner, double *z, const void *x, int64_t n, const MgPcgScalars *sc, hipStream_t s);
hipError_t mgpcg_pupdate(double *p, const double *z, int64_t n, const MgPcgScalars *sc, hipStream_t s);

// Host side of CGSolver (solvers_capi.cpp), shared by honei_hip_cg_solve and
// its allocation-free variant: argument checks (HONEI status code), the
// bytes of one CG vector (256-byte multiple), the matrix view, and the solve
// on caller-provided vectors p, r, v, scalars and reduction slots.
int cg_check_args(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows, size_t stride,
                  size_t threads, const void *b, const void *x);
size_t cg_vector_bytes(ValueType vt, size_t rows);
EllView cg_view(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows, size_t ncpr,
                size_t stride, size_t threads);
int cg_solve_on(const EllView &A, const void *b, void *x, size_t max_iters, double eps_relative, size_t *used_iters,
                double *final_defect, hipStream_t s, const ReduceSlots &slots, void *p, void *r, void *v,
                CGScalars *sc);
// Device address of the allocation-free variant's fixed scratch on the
// current device (looked up once per device); HONEI status code.
int noalloc_setup(void **out);
// Iteration `it` (0, 1, ... in order) of the loop.
hipError_t cg_iteration(const EllView &A, void *x, void *p, void *r, void *v, CGScalars *sc,
                        const ReduceSlots &sl, int it, hipStream_t s);
// Jacobi-preconditioned CG (math/conjugate_gradients.hh:884-913,
// 1110-1154).  pcg_start expects r = b - A x in r and ||r||^2 in
// sc->current; it forms d = dinv .* r and the scalars.  Vectors (and dinv)
// are double for F64 matrices, float otherwise.
hipError_t pcg_start(ValueType vt, void *d, const void *r, const void *dinv, int64_t n, CGScalars *sc,
                     double eps_rel, int max_iters, const ReduceSlots &sl, hipStream_t s);
hipError_t pcg_iteration(const EllView &A, void *x, void *d, void *r, void *t0, const void *dinv, CGScalars *sc,
                         const ReduceSlots &sl, int it, hipStream_t s);
hipError_t coarse_cg(const EllView &A, const void *rhs, void *x, void *g, void *u, void *en, double konv,
                     int max_iters, int *iters_dev, hipStream_t s);
hipError_t boundary_guess(double *guess, const double *b, int64_t n, int64_t side, hipStream_t s);
hipError_t acc_scaled_f32(double *result, const float *x, double scale, int64_t n, hipStream_t s);
hipError_t scale_to_f32(float *dst, const double *src, double inv, int64_t n, hipStream_t s);

// Device Q1 assembly (math/fill_matrix.hh:36-78 + la/sparse_matrix.hh:221-300
// + la/sparse_matrix_ell-impl.hh:75-122, T = 1): Dirichlet rows are identity,
// inner rows the 9-point stencil, entries in ascending column order.  With
// a per-cell coefficient (side-1)^2 array the stencil is the Q1 bilinear
// element stiffness summed over the four adjacent cells (coefficient 1
// reproduces 8/3 and -1/3).  dinv (optional) = omega / diag.
// threads > 1 writes the T-interleaved layout (stride >= side^2 * T); null Aj
// and Arl rewrite the values (and dinv) only.
hipError_t q1_assemble(ValueType vt, int64_t side, int32_t *Aj, void *Ax, uint32_t *Arl, int64_t stride,
                       const double *coeff, void *dinv, double omega, hipStream_t s, int threads = 1);

hipError_t random_ell(ValueType vt, int64_t n, int max_len, int64_t band, uint64_t seed, int32_t *Aj, void *Ax,
                      uint32_t *Arl, int64_t stride, int threads, hipStream_t s);
// Seeded fields: uniform in [lo, hi), and exp(sigma * N(0,1)) (Box-Muller).
hipError_t fill_uniform(double *dst, int64_t n, double lo, double hi, uint64_t seed, hipStream_t s);
hipError_t fill_lognormal(double *dst, int64_t n, double sigma, uint64_t seed, hipStream_t s);
// averaging: 0 arithmetic, 1 geometric, 2 harmonic mean of the 4 fine cells
hipError_t coarsen_coeff(const double *fine, int64_t nf_cells, double *coarse, int64_t nc_cells, int averaging,
                         hipStream_t s);

// ---------------------------------------------------------------------------
// Owned device ELL matrix
// ---------------------------------------------------------------------------
struct DevEll {
    EllView v{};
    int32_t *Aj = nullptr;
    void *Ax = nullptr;
    uint32_t *Arl = nullptr;
    bool owned = false;
    void release();
};

hipError_t dev_ell_alloc(DevEll &M, ValueType vt, int64_t rows, int64_t cols, int64_t stride, int64_t ncpr,
                         int threads);

size_t vt_size(ValueType vt);           // bytes per matrix value
size_t vec_size(ValueType vt);          // bytes per vector element (bf16 -> 4)

// ---------------------------------------------------------------------------
// Multigrid (math/multigrid.hh: MGInfo 66-149, _multigrid_kernel 1936-2250,
// FIXED driver 828-870, MIXED driver 2395-2536)
// ---------------------------------------------------------------------------
struct MGConfig {
    int64_t n_max_iter = 4, n_pre_smooth = 2, n_post_smooth = 2;
    int initial_zero = 0, convergence_check = 0, prolmat = 0, zero_inner_start = 0;
    double tolerance = 1e-8, adapt_correction_factor = 1.0;
    uint64_t mask[8] = {2, 2, 2, 2, 2, 2, 2, 2};
    int64_t max_outer = 0;          // 0 = unbounded (reference); capped at 10000
    int post_smooth_uses_pre = 1;   // reference: post-smoothing runs n_pre_smooth sweeps
};

struct MGLevel {
    int64_t n = 0, side = 0;
    DevEll A, P, R;                 // P: prolongation into this level; R: restriction from it
    void *dinv = nullptr;           // omega / diag(A)
    void *c = nullptr, *d = nullptr, *rhs = nullptr, *x = nullptr, *t = nullptr;
};

class MGHierarchy {
  public:
    MGHierarchy(ValueType inner, int64_t min_level, int64_t max_level);
    ~MGHierarchy();

    ValueType inner;                // matrix value type of the inner (V-cycle) levels
    int64_t min_level, max_level;
    std::vector<MGLevel> L;         // indexed by level
    DevEll A_outer;                 // FP64 system for MIXED (optional)
    MGConfig cfg;

    hipError_t alloc_level_vectors(int64_t level, int64_t n);
    hipError_t alloc_coarse_scratch();
    // Build every level (and the FP64 outer system) on the device as the
    // constant-coefficient Q1 Dirichlet problem on (2^l+1)^2 grids.
    hipError_t build_q1(double omega, hipStream_t s);
    // Same with a per-cell coefficient on the finest grid ((2^max-1)^2
    // doubles, device memory); coarse levels use the mean of the four fine
    // cells they cover (rediscretisation).
    hipError_t build_q1_coeff(double omega, const double *coeff_dev, int averaging, hipStream_t s);

    // One _multigrid_kernel call on the inner levels; returns the number of
    // V-cycles through *inner_iters.  rhs[max] and x[max] are inputs.
    hipError_t run_kernel(hipStream_t s, int64_t *inner_iters, double *final_defect);
    // Drivers; b and x are device vectors of the inner type (FIXED) or
    // double (MIXED).
    hipError_t solve_fixed(const void *b, void *x, hipStream_t s, int64_t *iters, double *final_defect);
    // FP64 conjugate gradients on the outer system preconditioned by one
    // lean inner V-cycle per iteration (z = |r| V(r/|r|)); start vector x
    // with the MIXED driver's boundary values; converged when the recursive
    // |r| < tol |r0| and a recomputed b - A x confirms it.
    hipError_t solve_pcg(const double *b, double *x, hipStream_t s, int64_t *iters, double *rel_defect);
    hipError_t solve_mixed(const double *b, double *x, hipStream_t s, int64_t *outer, int64_t *inner,
                           double *rel_defect, std::vector<double> *history);

  private:
    void *g_ = nullptr, *u_ = nullptr, *en_ = nullptr;   // coarse CG scratch
    int *iters_dev_ = nullptr;
    double *scal_ = nullptr;                            // device scalars
    double *guess_ = nullptr, *d64_ = nullptr, *res64_ = nullptr, *p64_ = nullptr, *q64_ = nullptr;
    int64_t n64_ = 0;
    bool mixed_inner_ = false;      // run_kernel called by the MIXED driver
    ReduceSlots slots_{};
    hipError_t smooth_from_zero(int64_t l, const void *rhs, void *out, void *acc, hipStream_t s);
    hipError_t sweeps(int64_t l, void *x, const void *rhs, int64_t count, hipStream_t s);
    hipError_t defect(int64_t l, void *out, const void *rhs, const void *x, double *norm2, hipStream_t s);
    hipError_t read_scalar(double *dst, const double *src, hipStream_t s);
};

}  // namespace hipx
}  // namespace honei
