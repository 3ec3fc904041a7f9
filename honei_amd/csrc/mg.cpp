// mg.cpp -- host-side multigrid drivers on device-resident hierarchies.
//
// Control flow follows the reference's monolithic V-cycle
// (math/multigrid.hh:1936-2250) and its FIXED (828-870) and MIXED
// (2395-2536) drivers step for step, including its quirks, so that a run
// can be compared with the CPU path iteration by iteration:
//   * pre-smoothing at the top level of a cycle smooths a correction c from
//     zero with the current defect as right-hand side, then x += c; on lower
//     levels x itself is smoothed from zero;
//   * post-smoothing runs n_pre_smooth sweeps (math/multigrid.hh:2168);
//   * the MIXED outer loop keeps the previous inner solution as the next
//     inner start vector unless cfg.zero_inner_start is set.
// Kernel fusion: the first two Jacobi sweeps from zero are one pass
// (ELL_JACOBI0) and the correction x += c is folded into the last sweep;
// prolongation and the coarse-grid correction x += alpha*c are one pass.

#include <cmath>
#include <cstring>

#include "runtime.hh"
#include "solvers.hh"

namespace honei {
namespace hipx {

size_t vt_size(ValueType vt) { return vt == VT_F64 ? 8 : (vt == VT_BF16 ? 2 : 4); }
size_t vec_size(ValueType vt) { return vt == VT_F64 ? 8 : 4; }

void DevEll::release() {
    if (owned) {
        if (Aj) (void)hipFree(Aj);
        if (Ax) (void)hipFree(Ax);
        if (Arl) (void)hipFree(Arl);
    }
    Aj = nullptr;
    Ax = nullptr;
    Arl = nullptr;
    owned = false;
}

hipError_t dev_ell_alloc(DevEll &M, ValueType vt, int64_t rows, int64_t cols, int64_t stride, int64_t ncpr,
                         int threads) {
    M.release();
    const size_t size = (size_t)stride * (size_t)(ncpr > 0 ? ncpr : 1);
    hipError_t e;
    if ((e = hipMalloc((void **)&M.Aj, size * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&M.Ax, size * vt_size(vt))) != hipSuccess) return e;
    if ((e = hipMalloc((void **)&M.Arl, (size_t)(rows > 0 ? rows : 1) * 4)) != hipSuccess) return e;
    M.owned = true;
    M.v.rows = rows;
    M.v.cols = cols;
    M.v.stride = stride;
    M.v.ncpr = ncpr;
    M.v.threads = threads;
    M.v.Aj = M.Aj;
    M.v.Ax = M.Ax;
    M.v.Arl = M.Arl;
    M.v.vtype = vt;
    return hipSuccess;
}

MGHierarchy::MGHierarchy(ValueType in, int64_t minl, int64_t maxl) : inner(in), min_level(minl), max_level(maxl) {
    L.resize((size_t)maxl + 1);
}

MGHierarchy::~MGHierarchy() {
    for (auto &l : L) {
        l.A.release();
        l.P.release();
        l.R.release();
        for (void *p : {l.dinv, l.c, l.d, l.rhs, l.x, l.t})
            if (p) (void)hipFree(p);
    }
    A_outer.release();
    for (void *p : {g_, u_, en_, (void *)iters_dev_, (void *)scal_, (void *)guess_, (void *)d64_, (void *)res64_,
                    (void *)p64_, (void *)q64_, (void *)slots_.partials, (void *)slots_.counter})
        if (p) (void)hipFree(p);
}

hipError_t MGHierarchy::alloc_level_vectors(int64_t level, int64_t n) {
    MGLevel &l = L[(size_t)level];
    l.n = n;
    int64_t s = 1;
    while (s * s < n) ++s;
    l.side = (s * s == n) ? s : 0;
    const size_t bytes = (size_t)(n > 0 ? n : 1) * vec_size(inner);
    for (void **p : {&l.dinv, &l.c, &l.d, &l.rhs, &l.x, &l.t}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        HX(hipMalloc(p, bytes));
        HX(fill_bytes_zero(*p, bytes, nullptr));
    }
    return hipSuccess;
}

hipError_t MGHierarchy::alloc_coarse_scratch() {
    const int64_t n = L[(size_t)min_level].n;
    const size_t bytes = (size_t)(n > 0 ? n : 1) * vec_size(inner);
    for (void **p : {&g_, &u_, &en_}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        HX(hipMalloc(p, bytes));
    }
    if (!iters_dev_) HX(hipMalloc((void **)&iters_dev_, 64));
    if (!scal_) {
        HX(hipMalloc((void **)&scal_, 64 * sizeof(double)));
        HX(fill_bytes_zero(scal_, 64 * sizeof(double), nullptr));
    }
    if (!slots_.partials) {
        slots_.capacity = 1 << 17;
        HX(hipMalloc((void **)&slots_.partials, sizeof(double) * (size_t)slots_.capacity * 4));
        HX(hipMalloc((void **)&slots_.counter, 64));
        HX(fill_bytes_zero(slots_.counter, 64, nullptr));
    }
    return hipDeviceSynchronize();
}

hipError_t MGHierarchy::build_q1(double omega, hipStream_t s) {
    for (int64_t l = min_level; l <= max_level; ++l) {
        const int64_t side = (1ll << l) + 1, n = side * side;
        const int64_t stride = 32 * ((n + 31) / 32);
        HX(alloc_level_vectors(l, n));
        HX(dev_ell_alloc(L[(size_t)l].A, inner, n, n, stride, side > 2 ? 9 : 1, 1));
        HX(q1_assemble(inner, side, L[(size_t)l].A.Aj, L[(size_t)l].A.Ax, L[(size_t)l].A.Arl, stride, nullptr,
                       L[(size_t)l].dinv, omega, s));
    }
    {
        const int64_t side = (1ll << max_level) + 1, n = side * side;
        const int64_t stride = 32 * ((n + 31) / 32);
        HX(dev_ell_alloc(A_outer, VT_F64, n, n, stride, 9, 1));
        HX(q1_assemble(VT_F64, side, A_outer.Aj, A_outer.Ax, A_outer.Arl, stride, nullptr, nullptr, 0.0, s));
    }
    HX(alloc_coarse_scratch());
    return hipStreamSynchronize(s);
}

hipError_t MGHierarchy::build_q1_coeff(double omega, const double *coeff, int averaging, hipStream_t s) {
    // per-level coefficient arrays, finest first
    std::vector<double *> kap((size_t)max_level + 1, nullptr);
    hipError_t err = hipSuccess;
    for (int64_t l = max_level; l >= min_level && err == hipSuccess; --l) {
        const int64_t cells = (1ll << l);
        if (l == max_level) {
            kap[(size_t)l] = (double *)coeff;
            continue;
        }
        err = hipMalloc((void **)&kap[(size_t)l], sizeof(double) * (size_t)(cells * cells));
        if (err == hipSuccess) err = coarsen_coeff(kap[(size_t)l + 1], 2 * cells, kap[(size_t)l], cells, averaging, s);
    }
    for (int64_t l = min_level; l <= max_level && err == hipSuccess; ++l) {
        const int64_t side = (1ll << l) + 1, n = side * side;
        const int64_t stride = 32 * ((n + 31) / 32);
        err = alloc_level_vectors(l, n);
        if (err == hipSuccess) err = dev_ell_alloc(L[(size_t)l].A, inner, n, n, stride, side > 2 ? 9 : 1, 1);
        if (err == hipSuccess)
            err = q1_assemble(inner, side, L[(size_t)l].A.Aj, L[(size_t)l].A.Ax, L[(size_t)l].A.Arl, stride,
                              kap[(size_t)l], L[(size_t)l].dinv, omega, s);
    }
    if (err == hipSuccess) {
        const int64_t side = (1ll << max_level) + 1, n = side * side;
        const int64_t stride = 32 * ((n + 31) / 32);
        err = dev_ell_alloc(A_outer, VT_F64, n, n, stride, 9, 1);
        if (err == hipSuccess)
            err = q1_assemble(VT_F64, side, A_outer.Aj, A_outer.Ax, A_outer.Arl, stride, coeff, nullptr, 0.0, s);
    }
    if (err == hipSuccess) err = alloc_coarse_scratch();
    (void)hipStreamSynchronize(s);
    for (int64_t l = min_level; l < max_level; ++l)
        if (kap[(size_t)l]) (void)hipFree(kap[(size_t)l]);
    return err;
}

hipError_t MGHierarchy::read_scalar(double *dst, const double *src, hipStream_t s) {
    HX(hipMemcpyAsync(dst, src, sizeof(double), hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

hipError_t MGHierarchy::defect(int64_t l, void *out, const void *rhs, const void *x, double *norm2, hipStream_t s) {
    MGLevel &lv = L[(size_t)l];
    EllArgs a;
    a.mode = ELL_DEFECT;
    a.out = out;
    a.x = x;
    a.rhs = rhs;
    a.norm2 = norm2;
    a.row_start = 0;
    a.row_end = lv.n;
    return ell_apply(lv.A.v, a, slots_, s);
}

// count Jacobi sweeps x <- x + dinv .* (rhs - A x), ping-ponging through t.
hipError_t MGHierarchy::sweeps(int64_t l, void *x, const void *rhs, int64_t count, hipStream_t s) {
    MGLevel &lv = L[(size_t)l];
    void *src = x, *dst = lv.t;
    for (int64_t k = 0; k < count; ++k) {
        EllArgs a;
        a.mode = ELL_JACOBI;
        a.out = dst;
        a.x = src;
        a.rhs = rhs;
        a.dinv = lv.dinv;
        a.row_start = 0;
        a.row_end = lv.n;
        HX(ell_apply(lv.A.v, a, slots_, s));
        void *tmp = src;
        src = dst;
        dst = tmp;
    }
    if (src != x) HX(vcopy(x, src, lv.n, inner == VT_F64 ? VT_F64 : VT_F32, s));
    return hipSuccess;
}

// out = n_pre Jacobi sweeps from zero with right-hand side rhs; acc += out
// when acc is non-null (Jacobi<Tag>::value first sweep + n-1 sweeps,
// math/jacobi.hh:481-513, then Sum).
hipError_t MGHierarchy::smooth_from_zero(int64_t l, const void *rhs, void *out, void *acc, hipStream_t s) {
    MGLevel &lv = L[(size_t)l];
    const ValueType vv = inner == VT_F64 ? VT_F64 : VT_F32;
    const int64_t n_pre = cfg.n_pre_smooth;
    if (n_pre <= 1) {
        HX(vprod(out, lv.dinv, rhs, lv.n, vv, s));
        if (acc) HX(vsum(acc, out, lv.n, vv, s));
        return hipSuccess;
    }
    // sweeps 1+2 fused; remaining sweeps ping-pong out <-> t
    const int64_t rest = n_pre - 2;
    void *first_dst = (rest % 2 == 0) ? out : lv.t;
    {
        EllArgs a;
        a.mode = ELL_JACOBI0;
        a.out = first_dst;
        a.rhs = rhs;
        a.dinv = lv.dinv;
        a.acc = rest == 0 ? acc : nullptr;
        a.row_start = 0;
        a.row_end = lv.n;
        HX(ell_apply(lv.A.v, a, slots_, s));
    }
    void *src = first_dst;
    for (int64_t k = 0; k < rest; ++k) {
        void *dst = (src == out) ? lv.t : out;
        EllArgs a;
        a.mode = ELL_JACOBI;
        a.out = dst;
        a.x = src;
        a.rhs = rhs;
        a.dinv = lv.dinv;
        a.acc = (k == rest - 1) ? acc : nullptr;
        a.row_start = 0;
        a.row_end = lv.n;
        HX(ell_apply(lv.A.v, a, slots_, s));
        src = dst;
    }
    return hipSuccess;
}

hipError_t MGHierarchy::run_kernel(hipStream_t s, int64_t *inner_iters, double *final_defect) {
    const int64_t maxl = max_level, minl = min_level;
    const ValueType vv = inner == VT_F64 ? VT_F64 : VT_F32;
    MGLevel &top = L[(size_t)maxl];
    double *nrm = scal_;   // scal_[0]: squared defect norm
    // Inside the MIXED driver the defect norms of the last allowed cycle are
    // never read (the driver forms its own FP64 defect), so that cycle skips
    // its final top-level defect and the host read; with one cycle per call
    // and no convergence test the initial norm is not needed either.  The
    // iterate x and the cycle count are unchanged by these omissions.
    const bool lean = mixed_inner_ && !cfg.convergence_check;
    const bool need_initial = !(lean && cfg.n_max_iter == 1);
    // d[max] = rhs when the start vector is zero (no matvec), else rhs - A x
    const void *top_d = top.d;
    if (cfg.initial_zero) {
        top_d = top.rhs;
        if (need_initial) HX(norm2(top.rhs, top.n, vv, nrm, slots_, s));
    } else {
        HX(defect(maxl, top.d, top.rhs, top.x, need_initial ? nrm : nullptr, s));
    }
    double h = 0.0;
    if (need_initial) HX(read_scalar(&h, nrm, s));
    // defect in the inner precision (Norm<l2,true,Tag> on Prec_)
    double defn = (inner == VT_F64) ? std::sqrt(h) : (double)std::sqrt((float)h);
    const double initial = defn;
    int64_t iters = 0;
    if (!(cfg.convergence_check && defn <= cfg.tolerance)) {
        int64_t iter;
        for (iter = 1; iter <= cfg.n_max_iter; ++iter) {
            const bool skip_final = lean && iter == cfg.n_max_iter;
            int64_t l = maxl;
            if (minl != maxl) {
                bool started = true;
                while (true) {
                    MGLevel &lv = L[(size_t)l];
                    if (started) {
                        HX(smooth_from_zero(l, (l == maxl) ? top_d : lv.d, lv.c, lv.x, s));
                    } else {
                        HX(smooth_from_zero(l, lv.rhs, lv.x, nullptr, s));
                    }
                    HX(defect(l, lv.d, lv.rhs, lv.x, nullptr, s));
                    --l;
                    MGLevel &lc = L[(size_t)l];
                    // d[l] = R d[l+1]; rhs[l] = d[l] (the restricted defect lives in rhs[l])
                    if (cfg.prolmat) {
                        EllArgs a;
                        a.mode = ELL_PRODUCT;
                        a.out = lc.rhs;
                        a.x = lv.d;
                        a.row_start = 0;
                        a.row_end = lc.n;
                        HX(ell_apply(lv.R.v, a, slots_, s));
                    } else {
                        HX(restriction(lc.rhs, lc.n, lv.d, lv.n, cfg.mask, vv, s));
                    }
                    if (l == minl) break;
                    started = false;
                }
            }
            {
                MGLevel &lm = L[(size_t)minl];
                const double konv = inner == VT_F64 ? 2.220446049250313e-16 : 1.1920928955078125e-07;
                const void *crhs = (minl == maxl) ? top_d : lm.rhs;
                HX(coarse_cg(lm.A.v, crhs, lm.x, g_, u_, en_, konv, (int)(4 * lm.n + 64), iters_dev_, s));
                if (minl == maxl) {
                    HX(defect(minl, lm.d, lm.rhs, lm.x, nrm, s));
                    break;
                }
            }
            while (true) {
                ++l;
                MGLevel &lv = L[(size_t)l];
                MGLevel &lc = L[(size_t)(l - 1)];
                if (cfg.adapt_correction_factor == 0.0) {
                    if (cfg.prolmat) {
                        EllArgs a;
                        a.mode = ELL_PRODUCT;
                        a.out = lv.c;
                        a.x = lc.x;
                        a.row_start = 0;
                        a.row_end = lv.n;
                        HX(ell_apply(lv.P.v, a, slots_, s));
                    } else {
                        HX(prolongation(lv.c, lv.n, lc.x, lc.n, cfg.mask, vv, s));
                    }
                    // alpha = (d, c) / (A c, c)
                    EllArgs a;
                    a.mode = ELL_PRODUCT;
                    a.out = lv.t;
                    a.x = lv.c;
                    a.row_start = 0;
                    a.row_end = lv.n;
                    HX(dot(lv.d, lv.c, lv.n, vv, scal_ + 1, slots_, s));
                    HX(ell_apply(lv.A.v, a, slots_, s));
                    HX(dot(lv.t, lv.c, lv.n, vv, scal_ + 2, slots_, s));
                    double d12[2];
                    HX(hipMemcpyAsync(d12, scal_ + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
                    HX(hipStreamSynchronize(s));
                    double alpha = (inner == VT_F64) ? d12[0] / d12[1] : (double)((float)d12[0] / (float)d12[1]);
                    HX(axpy(lv.x, lv.c, alpha, nullptr, lv.n, vv, s));
                } else if (cfg.prolmat) {
                    EllArgs a;
                    a.mode = ELL_PRODUCT;
                    a.out = lv.c;
                    a.x = lc.x;
                    a.row_start = 0;
                    a.row_end = lv.n;
                    HX(ell_apply(lv.P.v, a, slots_, s));
                    HX(axpy(lv.x, lv.c, cfg.adapt_correction_factor, nullptr, lv.n, vv, s));
                } else {
                    HX(prolongation_axpy(lv.c, lv.n, lc.x, lc.n, cfg.mask, lv.x, cfg.adapt_correction_factor, vv,
                                         s));
                }
                const int64_t post = cfg.post_smooth_uses_pre ? cfg.n_pre_smooth : cfg.n_post_smooth;
                HX(sweeps(l, lv.x, lv.rhs, post, s));
                // The reference also forms the defect after post-smoothing on
                // the lower levels; nothing reads it (the next down-leg
                // recomputes it first), so only the top level's is formed.
                if (l == maxl && !skip_final) HX(defect(l, lv.d, lv.rhs, lv.x, nrm, s));
                if (l == maxl) break;
            }
            if (skip_final) continue;   // loop ends: iter == n_max_iter
            HX(read_scalar(&h, nrm, s));
            defn = (inner == VT_F64) ? std::sqrt(h) : (double)std::sqrt((float)h);
            if (defn <= initial * cfg.tolerance) break;
        }
        iters = (iter > cfg.n_max_iter) ? iter - 1 : iter;
        if (minl == maxl) {
            HX(read_scalar(&h, nrm, s));
            defn = std::sqrt(h);
        }
    }
    if (inner_iters) *inner_iters = iters;
    if (final_defect) *final_defect = defn;
    return hipSuccess;
}

hipError_t MGHierarchy::solve_fixed(const void *b, void *x, hipStream_t s, int64_t *iters, double *final_defect) {
    MGLevel &top = L[(size_t)max_level];
    const ValueType vv = inner == VT_F64 ? VT_F64 : VT_F32;
    if (vv != VT_F64) return hipErrorInvalidValue;   // FIXED drivers run in FP64 here
    HX(boundary_guess((double *)x, (const double *)b, top.n, top.side, s));
    HX(vcopy(top.x, x, top.n, vv, s));
    HX(vcopy(top.rhs, b, top.n, vv, s));
    HX(run_kernel(s, iters, final_defect));
    HX(vcopy(x, top.x, top.n, vv, s));
    return hipStreamSynchronize(s);
}

// FP64 conjugate gradients preconditioned by one V-cycle of the inner
// hierarchy per iteration (z = |r| V(r / |r|), the cycle lean: from zero, no
// inner norms).  Every scalar stays on the device (MgPcgScalars): the host
// enqueues iterations in chunks, sized from the convergence rate seen so far,
// and reads the record once per chunk.  A residual that meets the tolerance
// is confirmed by the true defect b - A x before the solve stops, restarting
// CG from it when the two disagree.
hipError_t MGHierarchy::solve_pcg(const double *b, double *x, hipStream_t s, int64_t *iters_out,
                                  double *rel_defect) {
    MGLevel &top = L[(size_t)max_level];
    if (!A_outer.Aj) return hipErrorInvalidValue;
    const int64_t n = top.n;
    if (n64_ != n) {
        for (double **p : {&guess_, &d64_, &res64_, &p64_, &q64_}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
            HX(hipMalloc((void **)p, (size_t)n * sizeof(double)));
        }
        n64_ = n;
    }
    double *r = d64_, *z = res64_, *p = p64_, *q = q64_;
    static_assert(sizeof(MgPcgScalars) <= 32 * sizeof(double), "scalar record fits scal_[32..63]");
    MgPcgScalars *sc = (MgPcgScalars *)(scal_ + 32);
    const int cap = (int)(cfg.max_outer > 0 ? std::min<int64_t>(cfg.max_outer, 1 << 30) : 10000);
    // x0 with Dirichlet boundary values; r = b - A x0
    HX(boundary_guess(x, b, n, top.side, s));
    auto defect_norm = [&](double *norm2, const int *gate) -> hipError_t {
        EllArgs a;
        a.mode = ELL_DEFECT;
        a.out = r;
        a.x = x;
        a.rhs = b;
        a.norm2 = norm2;
        a.row_start = 0;
        a.row_end = n;
        a.gate = gate;
        return ell_apply(A_outer.v, a, slots_, s);
    };
    HX(defect_norm(&sc->r2, nullptr));
    HX(mgpcg_scalar(MGPCG_INIT, sc, cfg.tolerance, cap, s));
    // the inner cycle runs lean: one V-cycle from zero, no inner norms
    const MGConfig saved = cfg;
    cfg.n_max_iter = 1;
    cfg.initial_zero = 1;
    cfg.convergence_check = 0;
    auto iteration = [&]() -> hipError_t {
        HX(mgpcg_scale(inner, top.rhs, r, n, sc, s));
        HX(fill_bytes_zero(top.x, (size_t)n * vec_size(inner), s));
        int64_t inner_it = 0;
        mixed_inner_ = true;
        const hipError_t e = run_kernel(s, &inner_it, nullptr);
        mixed_inner_ = false;
        HX(e);
        HX(mgpcg_z(inner, z, top.x, n, sc, s));
        HX(dot(r, z, n, VT_F64, &sc->rz, slots_, s));
        HX(mgpcg_scalar(MGPCG_BETA, sc, 0.0, 0, s));
        HX(mgpcg_pupdate(p, z, n, sc, s));
        {
            EllArgs a;
            a.mode = ELL_PRODUCT;
            a.out = q;
            a.x = p;
            a.row_start = 0;
            a.row_end = n;
            HX(ell_apply(A_outer.v, a, slots_, s));
        }
        HX(dot(p, q, n, VT_F64, &sc->pq, slots_, s));
        HX(mgpcg_scalar(MGPCG_ALPHA, sc, 0.0, 0, s));
        HX(axpy(x, p, 0.0, &sc->alpha, n, VT_F64, s, &sc->done));
        HX(axpy_norm2(r, q, 0.0, &sc->malpha, n, VT_F64, &sc->r2, slots_, s, &sc->done));
        HX(mgpcg_scalar(MGPCG_CHECK, sc, 0.0, 0, s));
        HX(defect_norm(&sc->r2_true, &sc->verify));
        return mgpcg_scalar(MGPCG_CONFIRM, sc, 0.0, 0, s);
    };
    MgPcgScalars h{};
    hipError_t err = hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    int64_t chunk = 4;
    while (err == hipSuccess && !h.done) {
        for (int64_t k = 0; k < chunk && err == hipSuccess; ++k) err = iteration();
        if (err == hipSuccess) err = hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s);
        if (err == hipSuccess) err = hipStreamSynchronize(s);
        if (err != hipSuccess || h.done) break;
        // iterations still needed at the rate seen so far, 1..8 per chunk
        const double rel = h.initial > 0.0 ? h.nrm / h.initial : 1.0;
        int64_t need = 1;
        if (h.iters > 0 && rel > 0.0 && rel < 1.0 && std::isfinite(rel)) {
            const double per = std::log(rel) / (double)h.iters;
            need = (int64_t)std::ceil(std::log(cfg.tolerance / rel) / per);
        }
        chunk = std::max<int64_t>(1, std::min<int64_t>(8, need));
    }
    cfg = saved;
    HX(err);
    if (iters_out) *iters_out = h.iters;
    const double rd = h.initial > 0 ? h.nrm / h.initial : 0.0;
    if (rel_defect) *rel_defect = rd;
    return (h.initial > 0.0 && !(h.nrm < cfg.tolerance * h.initial)) ? hipErrorNotReady : hipSuccess;
}

hipError_t MGHierarchy::solve_mixed(const double *b, double *x, hipStream_t s, int64_t *outer_out,
                                    int64_t *inner_out, double *rel_defect, std::vector<double> *history) {
    MGLevel &top = L[(size_t)max_level];
    if (inner == VT_F64 || !A_outer.Aj) return hipErrorInvalidValue;
    const int64_t n = top.n;
    if (n64_ != n) {
        for (double **p : {&guess_, &d64_, &res64_, &p64_, &q64_}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
            HX(hipMalloc((void **)p, (size_t)n * sizeof(double)));
        }
        n64_ = n;
    }
    // x0 with Dirichlet boundary values, converted into the inner start vector
    HX(vcopy(guess_, x, n, VT_F64, s));
    HX(boundary_guess(guess_, b, n, top.side, s));
    HX(convert_d2f((float *)top.x, guess_, n, s));
    // D = b - A x0, d0 = ||D||
    double *nrm = scal_ + 8;
    {
        EllArgs a;
        a.mode = ELL_DEFECT;
        a.out = d64_;
        a.x = guess_;
        a.rhs = b;
        a.norm2 = nrm;
        a.row_start = 0;
        a.row_end = n;
        HX(ell_apply(A_outer.v, a, slots_, s));
    }
    double h;
    HX(read_scalar(&h, nrm, s));
    const double initial = std::sqrt(h);
    HX(fill_bytes_zero(res64_, (size_t)n * sizeof(double), s));
    double scale = 1.0, inv = 1.0;
    int64_t outer = 0, inner_total = 0;
    const int64_t cap = cfg.max_outer > 0 ? cfg.max_outer : 10000;
    double def_norm = initial;
    while (true) {
        HX(scale_to_f32((float *)top.rhs, d64_, inv, n, s));
        if (cfg.zero_inner_start) HX(fill_bytes_zero(top.x, (size_t)n * 4, s));
        int64_t it = 0;
        mixed_inner_ = true;
        hipError_t ek = run_kernel(s, &it, nullptr);
        mixed_inner_ = false;
        HX(ek);
        inner_total += it;
        ++outer;
        HX(acc_scaled_f32(res64_, (const float *)top.x, scale, n, s));
        EllArgs a;
        a.mode = ELL_DEFECT;
        a.out = d64_;
        a.x = res64_;
        a.rhs = b;
        a.norm2 = nrm;
        a.row_start = 0;
        a.row_end = n;
        HX(ell_apply(A_outer.v, a, slots_, s));
        HX(read_scalar(&h, nrm, s));
        def_norm = std::sqrt(h);
        if (history) history->push_back(initial > 0 ? def_norm / initial : 0.0);
        if (def_norm < cfg.tolerance * initial) break;
        if (outer >= cap || !std::isfinite(def_norm)) break;
        scale = def_norm;
        inv = 1.0 / scale;
    }
    HX(vcopy(x, res64_, n, VT_F64, s));
    HX(hipStreamSynchronize(s));
    if (outer_out) *outer_out = outer;
    if (inner_out) *inner_out = inner_total;
    if (rel_defect) *rel_defect = initial > 0 ? def_norm / initial : 0.0;
    return hipSuccess;
}

}  // namespace hipx
}  // namespace honei
