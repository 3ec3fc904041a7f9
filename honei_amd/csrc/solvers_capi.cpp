// solvers_capi.cpp -- extern "C" solver entry points (include/honei_hip.h).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "honei_hip.h"
#include "runtime.hh"
#include "solvers.hh"

using namespace honei::hipx;

namespace {

inline hipStream_t S(void *s) { return (hipStream_t)s; }

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

bool pow2_threads(size_t t) { return t == 1 || t == 2 || t == 4 || t == 8 || t == 16 || t == 32 || t == 64; }

// Upload a host ELL (uint64 indices, values of the given type) into an owned
// device matrix.  Values for a bf16 matrix are given as float.
int upload_ell(DevEll &M, ValueType vt, size_t rows, size_t cols, size_t stride, size_t ncpr, size_t threads,
               const uint64_t *Aj, const void *Ax, const uint64_t *Arl) {
    if (!Aj || !Ax || !Arl) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null ELL array");
    if (!pow2_threads(threads)) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "threads must be a power of two <= 64");
    if (rows * threads > stride) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "stride smaller than rows*threads");
    const size_t size = stride * ncpr;
    int st = check(dev_ell_alloc(M, vt, (int64_t)rows, (int64_t)cols, (int64_t)stride, (int64_t)ncpr, (int)threads),
                   "ELL allocation");
    if (st) return st;
    st = honei_hip_upload_indices(M.Aj, Aj, size, cols);
    if (st) return st;
    for (size_t r = 0; r < rows; ++r)
        if (Arl[r] > ncpr) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "row length exceeds num_cols_per_row");
    st = honei_hip_upload_indices(M.Arl, Arl, rows, UINT64_MAX);
    if (st) return st;
    if (vt == VT_BF16) return honei_hip_upload_values_bf16(M.Ax, (const float *)Ax, size);
    return check(hipMemcpy(M.Ax, Ax, size * vt_size(vt), hipMemcpyHostToDevice), "ELL value upload");
}

}  // namespace

namespace honei {
namespace hipx {

int cg_check_args(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows, size_t stride,
                  size_t threads, const void *b, const void *x) {
    if (vtype != HONEI_HIP_F32 && vtype != HONEI_HIP_F64 && vtype != HONEI_HIP_BF16)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!Aj || !Ax || !Arl || !b || !x) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    if (!pow2_threads(threads) || rows * threads > stride)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad ELL layout");
    if (rows > (size_t)INT32_MAX) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "rows exceed 32-bit index range");
    return HONEI_HIP_SUCCESS;
}

size_t cg_vector_bytes(ValueType vt, size_t rows) { return ((rows * vec_size(vt) + 255) / 256) * 256; }

EllView cg_view(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows, size_t ncpr,
                size_t stride, size_t threads) {
    EllView A;
    A.rows = (int64_t)rows;
    A.cols = (int64_t)rows;
    A.stride = (int64_t)stride;
    A.ncpr = (int64_t)ncpr;
    A.threads = (int)threads;
    A.Aj = (const int32_t *)Aj;
    A.Ax = Ax;
    A.Arl = (const uint32_t *)Arl;
    A.vtype = (ValueType)vtype;
    return A;
}

// CGSolver<Tag, NONE> (math/cg.hh:63-133) with p, r, v and the scalars in
// caller-provided device memory and the reductions in `slots`.
int cg_solve_on(const EllView &A, const void *b, void *x, size_t max_iters, double eps_relative, size_t *used_iters,
                double *final_defect, hipStream_t s, const ReduceSlots &slots, void *p, void *r, void *v,
                CGScalars *sc) {
    if (max_iters > 0x7fffffff) max_iters = 0x7fffffff;
    const size_t rows = (size_t)A.rows;
    const ValueType vv = A.vtype == VT_F64 ? VT_F64 : VT_F32;
    // r = b - A x ; p = r ; alpha = ||r||^2 (math/cg.hh:90-95)
    EllArgs a;
    a.mode = ELL_DEFECT;
    a.out = r;
    a.x = x;
    a.rhs = b;
    a.norm2 = &sc->alpha;
    a.row_start = 0;
    a.row_end = (int64_t)rows;
    int st = check(ell_apply(A, a, slots, s), "cg defect");
    if (st) return st;
    if ((st = check(vcopy(p, r, (int64_t)rows, vv, s), "cg copy"))) return st;
    if ((st = check(cg_init(vv, sc, eps_relative, (int)max_iters, s), "cg init"))) return st;
    // enqueue iterations in chunks; one flag read per chunk
    CGScalars h{};
    size_t chunk = rows < 100000 ? 16 : 4;
    size_t enqueued = 0;
    while (true) {
        for (size_t k = 0; k < chunk && enqueued < max_iters; ++k, ++enqueued)
            if ((st = check(cg_iteration(A, x, p, r, v, sc, slots, (int)enqueued, s), "cg iteration"))) return st;
        if ((st = check(hipMemcpyAsync(&h, sc, sizeof(CGScalars), hipMemcpyDeviceToHost, s), "cg readback")))
            return st;
        if ((st = check(hipStreamSynchronize(s), "cg sync"))) return st;
        if (h.done || enqueued >= max_iters) break;
        if (chunk < 64) chunk *= 2;
    }
    if (used_iters) *used_iters = (size_t)h.iters;
    if (final_defect) *final_defect = h.current;
    if (!h.done && max_iters > 0) return fail(HONEI_HIP_ERROR_NOT_CONVERGED, "CG did not converge");
    if (h.iters >= h.max_iters && !(h.current < h.eps_rel * h.initial || h.current < h.eps_rel))
        return fail(HONEI_HIP_ERROR_NOT_CONVERGED, "CG reached max_iters");
    return HONEI_HIP_SUCCESS;
}

}  // namespace hipx
}  // namespace honei

extern "C" {

// ---- CG -------------------------------------------------------------------
int honei_hip_cg_solve(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                       size_t num_cols_per_row, size_t stride, size_t threads, const void *b, void *x,
                       size_t max_iters, double eps_relative, size_t *used_iters, double *final_defect,
                       void *stream) {
    int st = cg_check_args(vtype, Aj, Ax, Arl, rows, stride, threads, b, x);
    if (st) return st;
    Workspace *w = workspace_for(S(stream));
    if (!w) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    const size_t vbytes = cg_vector_bytes((ValueType)vtype, rows);
    char *base = (char *)w->ensure_scratch(3 * vbytes + 256);
    if (!base) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    const EllView A = cg_view(vtype, Aj, Ax, Arl, rows, num_cols_per_row, stride, threads);
    return cg_solve_on(A, b, x, max_iters, eps_relative, used_iters, final_defect, S(stream), w->slots, base,
                       base + vbytes, base + 2 * vbytes, (CGScalars *)(base + 3 * vbytes));
}

int honei_hip_pcg_jacobi(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                         size_t num_cols_per_row, size_t stride, size_t threads, const void *b, void *x,
                         const void *dd_inverted, size_t max_iters, double eps_relative, size_t *used_iters,
                         double *final_defect, void *stream) {
    if (vtype != HONEI_HIP_F32 && vtype != HONEI_HIP_F64 && vtype != HONEI_HIP_BF16)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!Aj || !Ax || !Arl || !b || !x || !dd_inverted) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    if (!pow2_threads(threads) || rows * threads > stride)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad ELL layout");
    if (max_iters > 0x7fffffff) max_iters = 0x7fffffff;
    Workspace *w = workspace_for(S(stream));
    if (!w) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    const ValueType vt = (ValueType)vtype;
    const size_t es = vec_size(vt);
    const size_t vbytes = ((rows * es + 255) / 256) * 256;
    char *base = (char *)w->ensure_scratch(3 * vbytes + 256);
    if (!base) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    void *d = base, *r = base + vbytes, *t0 = base + 2 * vbytes;
    CGScalars *sc = (CGScalars *)(base + 3 * vbytes);
    EllView A;
    A.rows = (int64_t)rows;
    A.cols = (int64_t)rows;
    A.stride = (int64_t)stride;
    A.ncpr = (int64_t)num_cols_per_row;
    A.threads = (int)threads;
    A.Aj = (const int32_t *)Aj;
    A.Ax = Ax;
    A.Arl = (const uint32_t *)Arl;
    A.vtype = vt;
    hipStream_t s = S(stream);
    // r = b - A x ; ||r||^2 ; d = z = dinv .* r ; r.z
    // (math/conjugate_gradients.hh:1125-1134)
    EllArgs a;
    a.mode = ELL_DEFECT;
    a.out = r;
    a.x = x;
    a.rhs = b;
    a.norm2 = &sc->current;
    a.row_start = 0;
    a.row_end = (int64_t)rows;
    int st = check(ell_apply(A, a, w->slots, s), "pcg defect");
    if (st) return st;
    if ((st = check(pcg_start(vt, d, r, dd_inverted, (int64_t)rows, sc, eps_relative, (int)max_iters, w->slots, s),
                    "pcg start")))
        return st;
    // enqueue iterations in chunks; one flag read per chunk
    CGScalars h{};
    size_t chunk = rows < 100000 ? 16 : 4;
    size_t enqueued = 0;
    while (true) {
        for (size_t k = 0; k < chunk && enqueued < max_iters; ++k, ++enqueued)
            if ((st = check(pcg_iteration(A, x, d, r, t0, dd_inverted, sc, w->slots, (int)enqueued, s), "pcg iteration")))
                return st;
        if ((st = check(hipMemcpyAsync(&h, sc, sizeof(CGScalars), hipMemcpyDeviceToHost, s), "pcg readback")))
            return st;
        if ((st = check(hipStreamSynchronize(s), "pcg sync"))) return st;
        if (h.done || enqueued >= max_iters) break;
        if (chunk < 64) chunk *= 2;
    }
    if (used_iters) *used_iters = (size_t)h.iters;
    if (final_defect) *final_defect = h.current;
    if (max_iters > 0 && !h.converged) return fail(HONEI_HIP_ERROR_NOT_CONVERGED, "PCG-Jacobi reached max_iters");
    return HONEI_HIP_SUCCESS;
}

int honei_hip_cg_konv(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                      size_t num_cols_per_row, size_t stride, size_t threads, const void *b, void *x,
                      double konv_rad, size_t max_iters, size_t *iters, void *stream) {
    if (vtype != HONEI_HIP_F32 && vtype != HONEI_HIP_F64 && vtype != HONEI_HIP_BF16)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!Aj || !Ax || !Arl || !b || !x) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    if (!pow2_threads(threads) || rows * threads > stride)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad ELL layout");
    if (max_iters > 0x7fffffff) max_iters = 0x7fffffff;
    Workspace *w = workspace_for(S(stream));
    if (!w) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    const size_t es = vec_size((ValueType)vtype);
    const size_t vbytes = ((rows * es + 255) / 256) * 256;
    char *base = (char *)w->ensure_scratch(3 * vbytes + 256);
    if (!base) return HONEI_HIP_ERROR_OUT_OF_MEMORY;
    int *it_dev = (int *)(base + 3 * vbytes);
    EllView A;
    A.rows = (int64_t)rows;
    A.cols = (int64_t)rows;
    A.stride = (int64_t)stride;
    A.ncpr = (int64_t)num_cols_per_row;
    A.threads = (int)threads;
    A.Aj = (const int32_t *)Aj;
    A.Ax = Ax;
    A.Arl = (const uint32_t *)Arl;
    A.vtype = (ValueType)vtype;
    int st = check(coarse_cg(A, b, x, base, base + vbytes, base + 2 * vbytes, konv_rad, (int)max_iters, it_dev,
                             S(stream)),
                   "cg_konv");
    if (st) return st;
    int h = 0;
    if ((st = check(hipMemcpyAsync(&h, it_dev, sizeof(int), hipMemcpyDeviceToHost, S(stream)), "cg_konv readback")))
        return st;
    if ((st = check(hipStreamSynchronize(S(stream)), "cg_konv sync"))) return st;
    if (iters) *iters = (size_t)h;
    return HONEI_HIP_SUCCESS;
}

// ---- Q1 assembly -------------------------------------------------------------
int honei_hip_q1_assemble(int vtype, size_t side, void *Aj, void *Ax, void *Arl, size_t stride,
                          const double *coeff, void *dinv, double omega, size_t threads, void *stream) {
    if (vtype != HONEI_HIP_F32 && vtype != HONEI_HIP_F64 && vtype != HONEI_HIP_BF16)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!Aj || !Ax || !Arl) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    if (threads < 1 || threads > 32 || (threads & (threads - 1)))
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "threads must be a power of two <= 32");
    if (side < 2 || side > 46000 || stride < side * side * threads || stride % threads)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad side/stride");
    return check(q1_assemble((ValueType)vtype, (int64_t)side, (int32_t *)Aj, Ax, (uint32_t *)Arl,
                             (int64_t)stride, coeff, dinv, omega, S(stream), (int)threads),
                 "q1_assemble");
}

int honei_hip_random_ell(int vtype, size_t n, size_t max_len, size_t band, uint64_t seed, void *Aj, void *Ax,
                         void *Arl, size_t stride, size_t threads, void *stream) {
    if (vtype != HONEI_HIP_F32 && vtype != HONEI_HIP_F64) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (!Aj || !Ax || !Arl) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null pointer");
    if (n == 0 || n > (size_t)INT32_MAX || max_len < 1 || max_len > 64 || band < max_len || !pow2_threads(threads) ||
        stride < n * threads)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT,
                    "bad sizes (need 1<=max_len<=64, band>=max_len, threads a power of two <= 64, stride>=n*threads)");
    return check(random_ell((ValueType)vtype, (int64_t)n, (int)max_len, (int64_t)band, seed, (int32_t *)Aj, Ax,
                            (uint32_t *)Arl, (int64_t)stride, (int)threads, S(stream)),
                 "random_ell");
}

// ---- multigrid handle ------------------------------------------------------------
int honei_hip_mg_create(void **mg, int inner_vtype, size_t min_level, size_t max_level) {
    if (!mg) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "mg is null");
    if (inner_vtype != HONEI_HIP_F32 && inner_vtype != HONEI_HIP_F64 && inner_vtype != HONEI_HIP_BF16)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad vtype");
    if (min_level > max_level || max_level > 15) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "bad levels");
    MGHierarchy *h = new (std::nothrow) MGHierarchy((ValueType)inner_vtype, (int64_t)min_level, (int64_t)max_level);
    if (!h) return fail(HONEI_HIP_ERROR_OUT_OF_MEMORY, "host allocation failed");
    *mg = h;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_mg_destroy(void *mg) {
    (void)hipDeviceSynchronize();
    delete (MGHierarchy *)mg;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_mg_configure(void *mg, const honei_hip_mg_config *c) {
    if (!mg || !c) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    MGHierarchy *h = (MGHierarchy *)mg;
    if (c->n_pre_smooth == 0 || c->n_max_iter == 0) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "zero sweeps/cycles");
    h->cfg.n_max_iter = (int64_t)c->n_max_iter;
    h->cfg.n_pre_smooth = (int64_t)c->n_pre_smooth;
    h->cfg.n_post_smooth = (int64_t)c->n_post_smooth;
    h->cfg.initial_zero = c->initial_zero;
    h->cfg.convergence_check = c->convergence_check;
    h->cfg.prolmat = c->prolmat;
    h->cfg.zero_inner_start = c->zero_inner_start;
    h->cfg.tolerance = c->tolerance;
    h->cfg.adapt_correction_factor = c->adapt_correction_factor;
    for (int i = 0; i < 8; ++i) h->cfg.mask[i] = c->mask[i];
    h->cfg.max_outer = (int64_t)c->max_outer;
    h->cfg.post_smooth_uses_pre = c->post_smooth_uses_pre;
    return HONEI_HIP_SUCCESS;
}

int honei_hip_mg_set_level(void *mg, size_t level, size_t rows, size_t stride, size_t ncpr, size_t threads,
                           const uint64_t *Aj, const void *Ax, const uint64_t *Arl, const void *dinv) {
    if (!mg || !dinv) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    MGHierarchy *h = (MGHierarchy *)mg;
    if ((int64_t)level < h->min_level || (int64_t)level > h->max_level)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "level out of range");
    int st = upload_ell(h->L[level].A, h->inner, rows, rows, stride, ncpr, threads, Aj, Ax, Arl);
    if (st) return st;
    if ((st = check(h->alloc_level_vectors((int64_t)level, (int64_t)rows), "level vectors"))) return st;
    const size_t es = vec_size(h->inner);
    if ((st = check(hipMemcpy(h->L[level].dinv, dinv, rows * es, hipMemcpyHostToDevice), "dinv upload"))) return st;
    if ((int64_t)level == h->min_level) return check(h->alloc_coarse_scratch(), "coarse scratch");
    return HONEI_HIP_SUCCESS;
}

int honei_hip_mg_set_transfer(void *mg, size_t level, int kind, size_t rows, size_t cols, size_t stride,
                              size_t ncpr, size_t threads, const uint64_t *Aj, const void *Ax, const uint64_t *Arl) {
    if (!mg) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    MGHierarchy *h = (MGHierarchy *)mg;
    if ((int64_t)level <= h->min_level || (int64_t)level > h->max_level)
        return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "transfer level out of range");
    DevEll &M = kind == 0 ? h->L[level].P : h->L[level].R;
    return upload_ell(M, h->inner, rows, cols, stride, ncpr, threads, Aj, Ax, Arl);
}

int honei_hip_mg_set_outer(void *mg, size_t rows, size_t stride, size_t ncpr, size_t threads, const uint64_t *Aj,
                           const double *Ax, const uint64_t *Arl) {
    if (!mg) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    MGHierarchy *h = (MGHierarchy *)mg;
    return upload_ell(h->A_outer, VT_F64, rows, rows, stride, ncpr, threads, Aj, Ax, Arl);
}

int honei_hip_mg_build_q1(void *mg, double omega, void *stream) {
    if (!mg) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    return check(((MGHierarchy *)mg)->build_q1(omega, S(stream)), "mg_build_q1");
}

int honei_hip_mg_build_q1_coeff(void *mg, double omega, const double *coeff, int averaging, void *stream) {
    if (!mg || !coeff) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    if (averaging < 0 || averaging > 2) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "averaging must be 0, 1 or 2");
    return check(((MGHierarchy *)mg)->build_q1_coeff(omega, coeff, averaging, S(stream)), "mg_build_q1_coeff");
}

int honei_hip_mg_solve_fixed(void *mg, const void *b, void *x, size_t *iters, double *final_defect, void *stream) {
    if (!mg || !b || !x) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    MGHierarchy *h = (MGHierarchy *)mg;
    if (h->inner != VT_F64) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "FIXED solve needs an FP64 hierarchy");
    for (int64_t l = h->min_level; l <= h->max_level; ++l)
        if (!h->L[(size_t)l].A.Aj) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "hierarchy level missing");
    int64_t it = 0;
    double fd = 0;
    int st = check(h->solve_fixed(b, x, S(stream), &it, &fd), "mg_solve_fixed");
    if (iters) *iters = (size_t)it;
    if (final_defect) *final_defect = fd;
    return st;
}

int honei_hip_mg_solve_mixed(void *mg, const double *b, double *x, size_t *outer, size_t *inner,
                             double *rel_defect, void *stream) {
    if (!mg || !b || !x) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    MGHierarchy *h = (MGHierarchy *)mg;
    if (h->inner == VT_F64) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "MIXED solve needs an FP32/BF16 hierarchy");
    if (!h->A_outer.Aj) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "outer FP64 system not set");
    for (int64_t l = h->min_level; l <= h->max_level; ++l)
        if (!h->L[(size_t)l].A.Aj) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "hierarchy level missing");
    int64_t o = 0, in = 0;
    double rd = 0;
    int st = check(h->solve_mixed(b, x, S(stream), &o, &in, &rd, nullptr), "mg_solve_mixed");
    if (outer) *outer = (size_t)o;
    if (inner) *inner = (size_t)in;
    if (rel_defect) *rel_defect = rd;
    if (st) return st;
    if (!(rd < h->cfg.tolerance)) return fail(HONEI_HIP_ERROR_NOT_CONVERGED, "MIXED MG stopped before convergence");
    return HONEI_HIP_SUCCESS;
}

int honei_hip_mg_solve_pcg(void *mg, const double *b, double *x, size_t *iters, double *rel_defect, void *stream) {
    if (!mg || !b || !x) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    MGHierarchy *h = (MGHierarchy *)mg;
    if (!h->A_outer.Aj) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "outer FP64 system not set");
    for (int64_t l = h->min_level; l <= h->max_level; ++l)
        if (!h->L[(size_t)l].A.Aj) return fail(HONEI_HIP_ERROR_INVALID_ARGUMENT, "hierarchy level missing");
    int64_t it = 0;
    double rd = 0;
    const hipError_t e = h->solve_pcg(b, x, S(stream), &it, &rd);
    if (iters) *iters = (size_t)it;
    if (rel_defect) *rel_defect = rd;
    if (e == hipErrorNotReady) return fail(HONEI_HIP_ERROR_NOT_CONVERGED, "preconditioned CG stopped before convergence");
    return check(e, "mg_solve_pcg");
}

}  // extern "C"
