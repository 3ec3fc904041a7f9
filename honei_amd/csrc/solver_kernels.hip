// solver_kernels.hip -- device pieces of the CG and multigrid solvers.
//
// CG (math/cg.hh:63-133) runs as three kernels per iteration whose scalars
// (alpha, lambda, talpha, convergence flag, iteration count) live in device
// memory: the host never waits inside an iteration.  Kernels of an
// iteration that starts after convergence return at once, so the host may
// enqueue iterations in chunks and check the flag once per chunk.
//
// The coarse-grid CG of the multigrid V-cycle
// (math/conjugate_gradients.hh:154-201 driven by :587-623) runs inside one
// work-group: the coarsest Q1 grids have 9..1089 unknowns, far too small to
// amortise a launch per vector operation.

#include "solvers.hh"

#include "ell_device.hh"
#include "q1_device.hh"
#include "solver_device.hh"

#pragma clang fp contract(off)

namespace honei {
namespace hipx {
namespace {

// v = A p and partial v.p (one partial per block, summed by the next
// kernel).  Each block walks a contiguous range of 256/T-row chunks.  W:
// stored width q1_width(T), index and value streams read through buffer
// resources of aj_bytes / ax_bytes (the fixed-width row sum of the ELL row
// kernel; same sums).
template <typename VT, int T, bool W>
__global__ __launch_bounds__(256) void k_cg_spmv_dot(EllView A, const typename VecOf<VT>::type *p,
                                                     typename VecOf<VT>::type *v, CGScalars *sc, ReduceSlots sl,
                                                     int64_t nlogical, uint32_t aj_bytes, uint32_t ax_bytes) {
    typedef typename VecOf<VT>::type V;
    if (sc->done) return;
    const int64_t pb = xcd_block(blockIdx.x, gridDim.x);
    const int64_t per = nlogical / gridDim.x, rem = nlogical % gridDim.x;
    const int64_t first = pb * per + (pb < rem ? pb : rem);
    const int64_t last = first + per + (pb < rem ? 1 : 0);
    double part = 0.0;
    for (int64_t lb = first; lb < last; ++lb) {
        const int64_t gid = lb * (int64_t)blockDim.x + threadIdx.x;
        const int lane = (int)(gid & (T - 1));
        const int64_t row = gid / T;
        const bool active = row < A.rows;
        V s = V(0);
        if (active) {
            if constexpr (W)
                s = ell_lane_sum_buf<VT, T, q1_width(T)>(buffer_rsrc(A.Aj, aj_bytes), buffer_rsrc(A.Ax, ax_bytes), row,
                                                         lane, A.Arl[row], A.stride, p, nullptr);
            else
                s = ell_lane_sum<VT, T>(A.Aj, (const VT *)A.Ax, row, lane, A.Arl[row], A.stride, p);
        }
        s = lane_group_sum<T>(s);
        if (active && lane == 0) {
            v[row] = s;
            part += (double)s * (double)p[row];
        }
    }
    const double b = block_sum(part);
    if (threadIdx.x == 0) sl.partials[blockIdx.x] = b;
}

// lambda = alpha / (v.p) from the partials of k_cg_spmv_dot (one block)
template <typename V>
__global__ void k_cg_lambda(const double *partials, int64_t nparts, CGScalars *sc) {
    if (sc->done) return;
    const double tot = sum_partials(partials, nparts, 1, 0);
    if (threadIdx.x == 0) {
        const V temp = (V)tot;
        const V e = eps_of<V>();
        const V alpha = (V)sc->alpha;
        sc->lambda = (double)(alpha / ((temp < 0 ? -temp : temp) > e ? temp : e));
    }
}

// x += lambda p ; r -= lambda v ; alpha = ||r||^2 ; convergence test.
template <typename V>
__global__ __launch_bounds__(256) void k_cg_update(V *x, const V *p, V *r, const V *v, int64_t n, CGScalars *sc,
                                                   ReduceSlots sl) {
    if (sc->done) return;
    const V lambda = (V)sc->lambda;
    const V mlambda = -lambda;
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += p[i] * lambda;
        const V ri = r[i] + v[i] * mlambda;
        r[i] = ri;
        acc += (double)ri * (double)ri;
    }
    const double b = block_sum(acc);
    if (threadIdx.x == 0) sl.partials[blockIdx.x] = b;
}

// alpha = ||r||^2, talpha, iteration count and convergence test (one block)
template <typename V>
__global__ void k_cg_alpha(const double *partials, int64_t nparts, CGScalars *sc) {
    if (sc->done) return;
    const double tot = sum_partials(partials, nparts, 1, 0);
    if (threadIdx.x == 0) {
        const V e = eps_of<V>();
        const V alpha_old = (V)sc->alpha;
        const V alpha = (V)tot;
        sc->alpha = (double)alpha;
        sc->talpha = (double)(alpha / ((alpha_old < 0 ? -alpha_old : alpha_old) > e ? alpha_old : e));
        sc->iters += 1;
        const V current = sqrt(alpha);
        sc->current = (double)current;
        if (current < (V)sc->eps_rel * (V)sc->initial || current < (V)sc->eps_rel || sc->iters >= sc->max_iters)
            sc->done = 1;
    }
}

// p = p*talpha + r (Scale then Sum, math/cg.hh:119-120)
template <typename V>
__global__ __launch_bounds__(256) void k_cg_direction(V *p, const V *r, int64_t n, const CGScalars *sc) {
    if (sc->done) return;
    const V t = (V)sc->talpha;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const V q = p[i] * t;
        p[i] = q + r[i];
    }
}

template <typename V>
__global__ void k_cg_init(CGScalars *sc, double eps_rel, int max_iters) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const V alpha = (V)sc->alpha;
        sc->alpha = (double)alpha;
        sc->carry[0] = (double)alpha;
        sc->initial = (double)sqrt(alpha);
        sc->current = sc->initial;
        sc->eps_rel = eps_rel;
        sc->iters = 0;
        sc->max_iters = max_iters;
        sc->done = max_iters <= 0 ? 1 : 0;
    }
}

// ---- Jacobi-preconditioned CG ------------------------------------------------
// ConjugateGradients<Tag, JAC>::value on ELL (math/conjugate_gradients.hh:
// 1110-1154) with its cg_kernel (:884-913).  CGScalars carries alpha = r.z,
// lambda = the step (the reference's alpha), talpha = the direction factor
// (its beta), initial / current = ||r||.  z = dinv .* r is never stored: the
// update and direction kernels form it where they need it (same rounding).
// t0 = A d and d.t0 come from k_cg_spmv_dot (p = d, v = t0).

// d = dinv .* r and partial r.z (one block per slot of the partials)
template <typename V>
__global__ __launch_bounds__(256) void k_pcg_start(V *d, const V *r, const V *dinv, int64_t n, ReduceSlots sl) {
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const V ri = r[i];
        const V zi = dinv[i] * ri;
        d[i] = zi;
        acc += (double)ri * (double)zi;
    }
    const double b = block_sum(acc);
    if (threadIdx.x == 0) sl.partials[blockIdx.x] = b;
}

// r.z from the partials of k_pcg_start; ||r||^2 was left in sc->current by
// the defect kernel (one block)
template <typename V>
__global__ void k_pcg_init(const double *partials, int64_t nparts, CGScalars *sc, double eps_rel, int max_iters) {
    const double rz = sum_partials(partials, nparts, 1, 0);
    if (threadIdx.x == 0) {
        sc->alpha = (double)(V)rz;
        sc->carry[0] = (double)(V)rz;
        const V initial = sqrt((V)sc->current);
        sc->initial = (double)initial;
        sc->current = sc->initial;
        sc->eps_rel = eps_rel;
        sc->iters = 0;
        sc->max_iters = max_iters;
        sc->done = max_iters <= 0 ? 1 : 0;
        sc->converged = 0;
    }
}

// step = (r.z) / (d.t0), the denominator clamped to eps when |d.t0| < eps
// (one block)
template <typename V>
__global__ void k_pcg_step(const double *partials, int64_t nparts, CGScalars *sc) {
    if (sc->done) return;
    const double tot = sum_partials(partials, nparts, 1, 0);
    if (threadIdx.x == 0) {
        const V dt = (V)tot;
        const V e = eps_of<V>();
        const V rz = (V)sc->alpha;
        sc->lambda = (double)(((dt < 0 ? -dt : dt) >= e) ? rz / dt : rz / e);
    }
}

// x += step d ; r += -step t0 ; partials of z.r and r.r (z = dinv .* r)
template <typename V>
__global__ __launch_bounds__(256) void k_pcg_update(V *x, const V *d, V *r, const V *t0, const V *dinv, int64_t n,
                                                    CGScalars *sc, ReduceSlots sl) {
    if (sc->done) return;
    const V step = (V)sc->lambda;
    const V mstep = -step;
    double acc[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += d[i] * step;
        const V ri = r[i] + t0[i] * mstep;
        r[i] = ri;
        const V zi = dinv[i] * ri;
        acc[0] += (double)zi * (double)ri;
        acc[1] += (double)ri * (double)ri;
    }
    block_sum_first<2>(acc);
    if (threadIdx.x == 0) {
        sl.partials[2 * blockIdx.x] = acc[0];
        sl.partials[2 * blockIdx.x + 1] = acc[1];
    }
}

// direction factor = (z.r)_new / (r.z)_old (clamped like the step), ||r||,
// iteration count and the test ||r|| < initial * eps_rel (one block)
template <typename V>
__global__ void k_pcg_beta(const double *partials, int64_t nparts, CGScalars *sc) {
    if (sc->done) return;
    const double zr_t = sum_partials(partials, nparts, 2, 0);
    const double rr_t = sum_partials(partials, nparts, 2, 1);
    if (threadIdx.x == 0) {
        const V e = eps_of<V>();
        const V rz_old = (V)sc->alpha;
        const V zr = (V)zr_t;
        sc->talpha = (double)(((rz_old < 0 ? -rz_old : rz_old) >= e) ? zr / rz_old : zr / e);
        sc->alpha = (double)zr;
        sc->iters += 1;
        const V current = sqrt((V)rr_t);
        sc->current = (double)current;
        const bool conv = current < (V)sc->initial * (V)sc->eps_rel;
        sc->converged = conv ? 1 : 0;
        if (conv || sc->iters >= sc->max_iters) sc->done = 1;
    }
}

// d = z + d * factor with z = dinv .* r (ScaledSum(d, z, t0, beta) after
// copy(d, t0))
template <typename V>
__global__ __launch_bounds__(256) void k_pcg_direction(V *d, const V *r, const V *dinv, int64_t n,
                                                       const CGScalars *sc) {
    if (sc->done) return;
    const V t = (V)sc->talpha;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const V zi = dinv[i] * r[i];
        d[i] = zi + d[i] * t;
    }
}

// ---- three-kernel iteration for small systems --------------------------------
// When the SpMV and update launches have few blocks (<= kFewParts), the
// one-block scalar kernels are folded into their consumers: every block of
// the update sums the SpMV's partials itself (one block_sum, the same value
// in every block), and every block of the direction kernel sums the
// update's partials; block 0 of the direction kernel publishes the scalars.
// The carried scalar of iteration `it` is read from carry[it & 1] while the
// next one goes to carry[(it + 1) & 1].  Partials of the update live at
// kFusedOff so the SpMV of the next iteration cannot overwrite them early.
constexpr int kFewParts = 256;
constexpr int64_t kFusedOff = 8192;

template <typename V>
__global__ __launch_bounds__(256) void k_cg_update_f(V *x, const V *p, V *r, const V *v, int64_t n, CGScalars *sc,
                                                     ReduceSlots sl, int64_t nparts, int it) {
    if (sc->done) return;
    const double tot = sum_partials(sl.partials, nparts, 1, 0);
    const V temp = (V)tot;
    const V e = eps_of<V>();
    const V alpha = (V)sc->carry[it & 1];
    const V lambda = alpha / ((temp < 0 ? -temp : temp) > e ? temp : e);
    const V mlambda = -lambda;
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += p[i] * lambda;
        const V ri = r[i] + v[i] * mlambda;
        r[i] = ri;
        acc += (double)ri * (double)ri;
    }
    const double b = block_sum(acc);
    if (threadIdx.x == 0) sl.partials[kFusedOff + blockIdx.x] = b;
}

template <typename V>
__global__ __launch_bounds__(256) void k_cg_direction_f(V *p, const V *r, int64_t n, CGScalars *sc, ReduceSlots sl,
                                                        int64_t nparts, int it) {
    // block 0 of this launch publishes sc->done: read it once per block
    if (block_uniform_flag(&sc->done)) return;
    const double tot = sum_partials(sl.partials + kFusedOff, nparts, 1, 0);
    const V e = eps_of<V>();
    const V alpha_old = (V)sc->carry[it & 1];
    const V alpha = (V)tot;
    const V talpha = alpha / ((alpha_old < 0 ? -alpha_old : alpha_old) > e ? alpha_old : e);
    const V current = sqrt(alpha);
    const bool stop = current < (V)sc->eps_rel * (V)sc->initial || current < (V)sc->eps_rel || it + 1 >= sc->max_iters;
    if (!stop) {
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
            const V q = p[i] * talpha;
            p[i] = q + r[i];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->carry[(it + 1) & 1] = (double)alpha;
        sc->alpha = (double)alpha;
        sc->talpha = (double)talpha;
        sc->iters = it + 1;
        sc->current = (double)current;
        if (stop) sc->done = 1;
    }
}

template <typename V>
__global__ __launch_bounds__(256) void k_pcg_update_f(V *x, const V *d, V *r, const V *t0, const V *dinv, int64_t n,
                                                      CGScalars *sc, ReduceSlots sl, int64_t nparts, int it) {
    if (sc->done) return;
    const double tot = sum_partials(sl.partials, nparts, 1, 0);
    const V dt = (V)tot;
    const V e = eps_of<V>();
    const V rz = (V)sc->carry[it & 1];
    const V step = ((dt < 0 ? -dt : dt) >= e) ? rz / dt : rz / e;
    const V mstep = -step;
    double acc[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += d[i] * step;
        const V ri = r[i] + t0[i] * mstep;
        r[i] = ri;
        const V zi = dinv[i] * ri;
        acc[0] += (double)zi * (double)ri;
        acc[1] += (double)ri * (double)ri;
    }
    block_sum_first<2>(acc);
    if (threadIdx.x == 0) {
        sl.partials[kFusedOff + 2 * blockIdx.x] = acc[0];
        sl.partials[kFusedOff + 2 * blockIdx.x + 1] = acc[1];
    }
}

template <typename V>
__global__ __launch_bounds__(256) void k_pcg_direction_f(V *d, const V *r, const V *dinv, int64_t n, CGScalars *sc,
                                                         ReduceSlots sl, int64_t nparts, int it) {
    // block 0 of this launch publishes sc->done: read it once per block
    if (block_uniform_flag(&sc->done)) return;
    const double zr_t = sum_partials(sl.partials + kFusedOff, nparts, 2, 0);
    const double rr_t = sum_partials(sl.partials + kFusedOff, nparts, 2, 1);
    const V e = eps_of<V>();
    const V rz_old = (V)sc->carry[it & 1];
    const V zr = (V)zr_t;
    const V factor = ((rz_old < 0 ? -rz_old : rz_old) >= e) ? zr / rz_old : zr / e;
    const V current = sqrt((V)rr_t);
    const bool conv = current < (V)sc->initial * (V)sc->eps_rel;
    const bool stop = conv || it + 1 >= sc->max_iters;
    if (!stop) {
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
            const V zi = dinv[i] * r[i];
            d[i] = zi + d[i] * factor;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->carry[(it + 1) & 1] = (double)zr;
        sc->alpha = (double)zr;
        sc->talpha = (double)factor;
        sc->iters = it + 1;
        sc->current = (double)current;
        sc->converged = conv ? 1 : 0;
        if (stop) sc->done = 1;
    }
}

template <typename VT>
__global__ __launch_bounds__(256) void k_coarse_cg(EllView A, const typename VecOf<VT>::type *rhs,
                                                   typename VecOf<VT>::type *x, typename VecOf<VT>::type *g,
                                                   typename VecOf<VT>::type *u, typename VecOf<VT>::type *en,
                                                   double konv, int max_iters, int *iters_out) {
    const int it = coarse_cg_block<VT>(A, rhs, x, g, u, en, konv, max_iters);
    if (threadIdx.x == 0 && iters_out) *iters_out = it;
}

// Multigrid helpers ---------------------------------------------------------
__global__ void k_boundary_guess(double *guess, const double *b, int64_t n, int64_t side) {
    // first grid row, last grid row, first grid column take b
    // (math/multigrid.hh:2423-2440)
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (guess_node(i, n, side)) guess[i] = b[i];
}

__global__ void k_acc_scaled_f32(double *result, const float *x, double scale, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        result[i] += (double)x[i] * scale;
}

__global__ void k_scale_to_f32(float *dst, const double *src, double inv, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)(src[i] * inv);
}

// ---- multigrid-preconditioned CG with device scalars (MGHierarchy::solve_pcg) --------
// One thread updates the record; the vector kernels read it once per block.
__global__ void k_mgpcg_scalar(int step, MgPcgScalars *sc, double tol, int cap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    MgPcgScalars &c = *sc;
    if (step == MGPCG_INIT) {   // after r = b - A x0 with ||r||^2 in r2
        c.initial = sqrt(c.r2);
        c.nrm = c.initial;
        c.rho_prev = 0.0;
        c.tol = tol;
        c.iters = 0;
        c.cap = cap;
        c.verify = 0;
        c.alpha = c.malpha = c.beta = 0.0;
        c.done = (c.initial > 0.0 && cap > 0 && isfinite(c.nrm)) ? 0 : 1;
        return;
    }
    if (c.done) return;
    if (step == MGPCG_BETA) {   // after r.z
        c.iters += 1;
        c.beta = (c.rho_prev > 0.0 && c.rz > 0.0) ? c.rz / c.rho_prev : 0.0;
    } else if (step == MGPCG_ALPHA) {   // after p.q
        if (!(c.pq > 0.0) || !(c.rz > 0.0)) {
            c.done = 1;   // breakdown: not converged
            return;
        }
        c.alpha = c.rz / c.pq;
        c.malpha = -c.alpha;
    } else if (step == MGPCG_CHECK) {   // after x += alpha p, r -= alpha q, ||r||^2
        c.nrm = sqrt(c.r2);
        c.rho_prev = c.rz;
        if (c.nrm < c.tol * c.initial)
            c.verify = 1;
        else if (!(c.iters < c.cap) || !isfinite(c.nrm))
            c.done = 1;
    } else if (step == MGPCG_CONFIRM) {   // after the gated r = b - A x, ||r||^2 in r2_true
        if (!c.verify) return;
        c.verify = 0;
        c.nrm = sqrt(c.r2_true);
        if (c.nrm < c.tol * c.initial) {
            c.done = 1;
        } else {
            c.rho_prev = 0.0;   // drifted: restart from the true residual
            if (!(c.iters < c.cap) || !isfinite(c.nrm)) c.done = 1;
        }
    }
}

// rhs = r / |r| in the inner precision (the rounding of r * (1 / |r|))
template <typename V>
__global__ void k_mgpcg_scale(V *rhs, const double *r, int64_t n, const MgPcgScalars *sc) {
    if (block_uniform_flag(&sc->done)) return;
    const double inv = 1.0 / sc->nrm;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        rhs[i] = (V)(r[i] * inv);
}

// z = |r| x (FP32 inner solution: widened and scaled into a zero vector)
template <typename V>
__global__ void k_mgpcg_z(double *z, const V *x, int64_t n, const MgPcgScalars *sc) {
    if (block_uniform_flag(&sc->done)) return;
    const double nrm = sc->nrm;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (sizeof(V) == 4)
            z[i] = 0.0 + (double)x[i] * nrm;
        else
            z[i] = x[i] * nrm;
    }
}

// p = z (beta 0) or z + p beta
__global__ void k_mgpcg_pupdate(double *p, const double *z, int64_t n, const MgPcgScalars *sc) {
    if (block_uniform_flag(&sc->done)) return;
    const double beta = sc->beta;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = beta == 0.0 ? z[i] : z[i] + p[i] * beta;
}

inline int grid_for(int64_t n, int per = 1024) {
    int64_t g = (n + per - 1) / per;
    return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

#define HX_RET(expr)                          \
    do {                                      \
        const hipError_t e_ = (expr);         \
        if (e_ != hipSuccess) return e_;      \
    } while (0)

// t = A p with the partials of t.p; the fixed-width form where the stored
// width is that of the Q1 operator and every byte offset fits 31 bits
template <typename VT>
hipError_t launch_spmv_dot(const EllView &A, const void *p, void *t, CGScalars *sc, const ReduceSlots &sl,
                           int64_t nlogical, int64_t grid, hipStream_t s) {
    typedef typename VecOf<VT>::type V;
    const int T = A.threads;
    const int64_t slots_total = A.ncpr * A.stride;
    const bool wide = T <= 8 && A.ncpr == q1_width(T) && slots_total * 8 < (int64_t(1) << 31);
    const uint32_t ajb = (uint32_t)(wide ? slots_total * 4 : 0);
    const uint32_t axb = (uint32_t)(wide ? slots_total * (int64_t)sizeof(VT) : 0);
    const dim3 g((unsigned)grid), b(256);
#define SDK(TT)                                                                                                  \
    case TT:                                                                                                     \
        if (wide && TT <= 8)                                                                                     \
            hipLaunchKernelGGL((k_cg_spmv_dot<VT, TT, (TT <= 8)>), g, b, 0, s, A, (const V *)p, (V *)t, sc, sl,  \
                               nlogical, ajb, axb);                                                              \
        else                                                                                                     \
            hipLaunchKernelGGL((k_cg_spmv_dot<VT, TT, false>), g, b, 0, s, A, (const V *)p, (V *)t, sc, sl,      \
                               nlogical, ajb, axb);                                                              \
        break;
    switch (T) {
        SDK(1) SDK(2) SDK(4) SDK(8) SDK(16) SDK(32) SDK(64)
        default: return hipErrorInvalidValue;
    }
#undef SDK
    return hipGetLastError();
}

template <typename VT>
hipError_t cg_iter_t(const EllView &A, void *x, void *p, void *r, void *v, CGScalars *sc, const ReduceSlots &sl,
                     int it, hipStream_t s) {
    typedef typename VecOf<VT>::type V;
    const int T = A.threads;
    const int64_t nlogical = (A.rows * T + 255) / 256;
    const int64_t grid = nlogical < sl.capacity ? nlogical : sl.capacity;
    HX_RET(launch_spmv_dot<VT>(A, p, v, sc, sl, nlogical, grid, s));
    const int g2 = grid_for(A.rows, 2048);
    if (grid <= kFewParts && g2 <= kFewParts) {
        hipLaunchKernelGGL((k_cg_update_f<V>), dim3(g2), dim3(256), 0, s, (V *)x, (const V *)p, (V *)r, (const V *)v,
                           A.rows, sc, sl, grid, it);
        hipLaunchKernelGGL((k_cg_direction_f<V>), dim3(grid_for(A.rows)), dim3(256), 0, s, (V *)p, (const V *)r,
                           A.rows, sc, sl, (int64_t)g2, it);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_cg_lambda<V>), dim3(1), dim3(1024), 0, s, (const double *)sl.partials, grid, sc);
    hipLaunchKernelGGL((k_cg_update<V>), dim3(g2), dim3(256), 0, s, (V *)x, (const V *)p, (V *)r, (const V *)v,
                       A.rows, sc, sl);
    hipLaunchKernelGGL((k_cg_alpha<V>), dim3(1), dim3(1024), 0, s, (const double *)sl.partials, (int64_t)g2, sc);
    hipLaunchKernelGGL((k_cg_direction<V>), dim3(grid_for(A.rows)), dim3(256), 0, s, (V *)p, (const V *)r, A.rows,
                       (const CGScalars *)sc);
    return hipGetLastError();
}

template <typename VT>
hipError_t pcg_iter_t(const EllView &A, void *x, void *d, void *r, void *t0, const void *dinv, CGScalars *sc,
                      const ReduceSlots &sl, int it, hipStream_t s) {
    typedef typename VecOf<VT>::type V;
    const int T = A.threads;
    const int64_t nlogical = (A.rows * T + 255) / 256;
    const int64_t grid = nlogical < sl.capacity ? nlogical : sl.capacity;
    HX_RET(launch_spmv_dot<VT>(A, d, t0, sc, sl, nlogical, grid, s));
    const int g2 = grid_for(A.rows, 2048);
    if (grid <= kFewParts && g2 <= kFewParts) {
        hipLaunchKernelGGL((k_pcg_update_f<V>), dim3(g2), dim3(256), 0, s, (V *)x, (const V *)d, (V *)r,
                           (const V *)t0, (const V *)dinv, A.rows, sc, sl, grid, it);
        hipLaunchKernelGGL((k_pcg_direction_f<V>), dim3(grid_for(A.rows)), dim3(256), 0, s, (V *)d, (const V *)r,
                           (const V *)dinv, A.rows, sc, sl, (int64_t)g2, it);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_pcg_step<V>), dim3(1), dim3(1024), 0, s, (const double *)sl.partials, grid, sc);
    hipLaunchKernelGGL((k_pcg_update<V>), dim3(g2), dim3(256), 0, s, (V *)x, (const V *)d, (V *)r, (const V *)t0,
                       (const V *)dinv, A.rows, sc, sl);
    hipLaunchKernelGGL((k_pcg_beta<V>), dim3(1), dim3(1024), 0, s, (const double *)sl.partials, (int64_t)g2, sc);
    hipLaunchKernelGGL((k_pcg_direction<V>), dim3(grid_for(A.rows)), dim3(256), 0, s, (V *)d, (const V *)r,
                       (const V *)dinv, A.rows, (const CGScalars *)sc);
    return hipGetLastError();
}

template <typename V>
hipError_t pcg_start_t(void *d, const void *r, const void *dinv, int64_t n, CGScalars *sc, double eps_rel,
                       int max_iters, const ReduceSlots &sl, hipStream_t s) {
    const int g = grid_for(n, 2048);
    hipLaunchKernelGGL((k_pcg_start<V>), dim3(g), dim3(256), 0, s, (V *)d, (const V *)r, (const V *)dinv, n, sl);
    hipLaunchKernelGGL((k_pcg_init<V>), dim3(1), dim3(1024), 0, s, (const double *)sl.partials, (int64_t)g, sc,
                       eps_rel, max_iters);
    return hipGetLastError();
}

}  // namespace

hipError_t cg_iteration(const EllView &A, void *x, void *p, void *r, void *v, CGScalars *sc,
                        const ReduceSlots &sl, int it, hipStream_t s) {
    switch (A.vtype) {
        case VT_F32: return cg_iter_t<float>(A, x, p, r, v, sc, sl, it, s);
        case VT_F64: return cg_iter_t<double>(A, x, p, r, v, sc, sl, it, s);
        case VT_BF16: return cg_iter_t<bf16>(A, x, p, r, v, sc, sl, it, s);
    }
    return hipErrorInvalidValue;
}

hipError_t pcg_iteration(const EllView &A, void *x, void *d, void *r, void *t0, const void *dinv, CGScalars *sc,
                         const ReduceSlots &sl, int it, hipStream_t s) {
    switch (A.vtype) {
        case VT_F32: return pcg_iter_t<float>(A, x, d, r, t0, dinv, sc, sl, it, s);
        case VT_F64: return pcg_iter_t<double>(A, x, d, r, t0, dinv, sc, sl, it, s);
        case VT_BF16: return pcg_iter_t<bf16>(A, x, d, r, t0, dinv, sc, sl, it, s);
    }
    return hipErrorInvalidValue;
}

hipError_t pcg_start(ValueType vt, void *d, const void *r, const void *dinv, int64_t n, CGScalars *sc,
                     double eps_rel, int max_iters, const ReduceSlots &sl, hipStream_t s) {
    if (vt == VT_F64) return pcg_start_t<double>(d, r, dinv, n, sc, eps_rel, max_iters, sl, s);
    return pcg_start_t<float>(d, r, dinv, n, sc, eps_rel, max_iters, sl, s);
}

hipError_t cg_init(ValueType vt, CGScalars *sc, double eps_rel, int max_iters, hipStream_t s) {
    if (vt == VT_F64)
        hipLaunchKernelGGL(k_cg_init<double>, dim3(1), dim3(64), 0, s, sc, eps_rel, max_iters);
    else
        hipLaunchKernelGGL(k_cg_init<float>, dim3(1), dim3(64), 0, s, sc, eps_rel, max_iters);
    return hipGetLastError();
}

hipError_t coarse_cg(const EllView &A, const void *rhs, void *x, void *g, void *u, void *en, double konv,
                     int max_iters, int *iters_dev, hipStream_t s) {
    switch (A.vtype) {
        case VT_F32:
            hipLaunchKernelGGL(k_coarse_cg<float>, dim3(1), dim3(256), 0, s, A, (const float *)rhs, (float *)x,
                               (float *)g, (float *)u, (float *)en, konv, max_iters, iters_dev);
            break;
        case VT_F64:
            hipLaunchKernelGGL(k_coarse_cg<double>, dim3(1), dim3(256), 0, s, A, (const double *)rhs, (double *)x,
                               (double *)g, (double *)u, (double *)en, konv, max_iters, iters_dev);
            break;
        case VT_BF16:
            hipLaunchKernelGGL(k_coarse_cg<bf16>, dim3(1), dim3(256), 0, s, A, (const float *)rhs, (float *)x,
                               (float *)g, (float *)u, (float *)en, konv, max_iters, iters_dev);
            break;
    }
    return hipGetLastError();
}

hipError_t boundary_guess(double *guess, const double *b, int64_t n, int64_t side, hipStream_t s) {
    hipLaunchKernelGGL(k_boundary_guess, dim3(grid_for(n)), dim3(256), 0, s, guess, b, n, side);
    return hipGetLastError();
}
hipError_t acc_scaled_f32(double *result, const float *x, double scale, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_acc_scaled_f32, dim3(grid_for(n)), dim3(256), 0, s, result, x, scale, n);
    return hipGetLastError();
}
hipError_t mgpcg_scalar(int step, MgPcgScalars *sc, double tol, int cap, hipStream_t s) {
    hipLaunchKernelGGL(k_mgpcg_scalar, dim3(1), dim3(64), 0, s, step, sc, tol, cap);
    return hipGetLastError();
}
hipError_t mgpcg_scale(ValueType inner, void *rhs, const double *r, int64_t n, const MgPcgScalars *sc,
                       hipStream_t s) {
    if (inner == VT_F64)
        hipLaunchKernelGGL(k_mgpcg_scale<double>, dim3(grid_for(n)), dim3(256), 0, s, (double *)rhs, r, n, sc);
    else
        hipLaunchKernelGGL(k_mgpcg_scale<float>, dim3(grid_for(n)), dim3(256), 0, s, (float *)rhs, r, n, sc);
    return hipGetLastError();
}
hipError_t mgpcg_z(ValueType inner, double *z, const void *x, int64_t n, const MgPcgScalars *sc, hipStream_t s) {
    if (inner == VT_F64)
        hipLaunchKernelGGL(k_mgpcg_z<double>, dim3(grid_for(n)), dim3(256), 0, s, z, (const double *)x, n, sc);
    else
        hipLaunchKernelGGL(k_mgpcg_z<float>, dim3(grid_for(n)), dim3(256), 0, s, z, (const float *)x, n, sc);
    return hipGetLastError();
}
hipError_t mgpcg_pupdate(double *p, const double *z, int64_t n, const MgPcgScalars *sc, hipStream_t s) {
    hipLaunchKernelGGL(k_mgpcg_pupdate, dim3(grid_for(n)), dim3(256), 0, s, p, z, n, sc);
    return hipGetLastError();
}
hipError_t scale_to_f32(float *dst, const double *src, double inv, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_to_f32, dim3(grid_for(n)), dim3(256), 0, s, dst, src, inv, n);
    return hipGetLastError();
}

}  // namespace hipx
}  // namespace honei
