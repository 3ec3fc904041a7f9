// honei/math/multigrid.hh -- MGInfo and Multigrid<tags::GPU::HIP, ...>
// (math/multigrid.hh:66-149, 828-870, 2395-2536).
//
// The hierarchy in MGInfo (matrices, inverted diagonals, transfer matrices)
// is uploaded once into a device-resident hierarchy (honei_hip_mg_*) and
// cached in the MGInfo; value() then runs the whole FIXED or MIXED solve on
// the device with the reference's control flow.
#pragma once

#include <cmath>
#include <memory>
#include <vector>

#include "honei/backends/hip/device.hh"

namespace honei {

namespace hip {
struct MGHandle {
    void *h = nullptr;
    ~MGHandle() {
        if (h) honei_hip_mg_destroy(h);
    }
};
}  // namespace hip

template <typename Prec_, typename DataLayout_>
struct MGInfo {
    bool is_smoother = false;
    DenseVector<unsigned long> macro_border_mask;
    unsigned long min_level = 1, max_level = 1;
    unsigned long n_max_iter = 4;
    bool initial_zero = false;
    Prec_ tolerance = Prec_(1e-8);
    bool convergence_check = false;
    unsigned long inner_iterations = 0;
    unsigned long n_pre_smooth = 2, n_post_smooth = 2, n_max_iter_coarse = 100;
    Prec_ tolerance_coarse = Prec_(1e-2);
    Prec_ adapt_correction_factor = Prec_(1);

    std::vector<DenseVector<Prec_>> c, d, rhs, x, diags_inverted, temp_0, temp_1;
    std::vector<DataLayout_> a, prolmats, resmats, spais;

    // Extensions of this backend (not in the reference): start every outer
    // step of the MIXED driver from a zero inner iterate, a cap on outer
    // steps, outer iteration count of the last solve, and the device copy.
    bool zero_inner_start = false;
    unsigned long max_outer = 0;
    unsigned long outer_iterations = 0;
    std::shared_ptr<hip::MGHandle> device;

    MGInfo() : macro_border_mask(8, 2ul) {}
};

template <typename Tag_, typename OuterTag_, typename ProlType_, typename Smoother_, typename Cycle_, typename Mode_>
struct Multigrid;

namespace hip {

template <typename Prec_>
void upload_ell(void *mg, unsigned long level, const SparseMatrixELL<Prec_> &A, int kind, bool as_level,
                const DenseVector<Prec_> *dinv) {
    const uint64_t *aj = (const uint64_t *)A.Aj().elements();
    const uint64_t *arl = (const uint64_t *)A.Arl().elements();
    const void *ax = A.Ax().elements();
    if (as_level)
        hip_check(honei_hip_mg_set_level(mg, level, A.rows(), A.stride(), A.num_cols_per_row(), A.threads(), aj, ax,
                                         arl, dinv->elements()),
                  "mg_set_level");
    else
        hip_check(honei_hip_mg_set_transfer(mg, level, kind, A.rows(), A.columns(), A.stride(), A.num_cols_per_row(),
                                            A.threads(), aj, ax, arl),
                  "mg_set_transfer");
}

template <typename Prec_, typename Prol_>
void *ensure_hierarchy(MGInfo<Prec_, SparseMatrixELL<Prec_>> &info, const SparseMatrixELL<double> *outer) {
    if (!info.device) {
        auto h = std::make_shared<MGHandle>();
        hip_check(honei_hip_mg_create(&h->h, HipType<Prec_>::value, info.min_level, info.max_level), "mg_create");
        for (unsigned long l = info.min_level; l <= info.max_level; ++l) {
            upload_ell<Prec_>(h->h, l, info.a[l], 0, true, &info.diags_inverted[l]);
            if (std::is_same<Prol_, methods::PROLMAT>::value && l > info.min_level) {
                upload_ell<Prec_>(h->h, l, info.prolmats[l], 0, false, nullptr);
                upload_ell<Prec_>(h->h, l, info.resmats[l], 1, false, nullptr);
            }
        }
        if (outer) {
            hip_check(honei_hip_mg_set_outer(h->h, outer->rows(), outer->stride(), outer->num_cols_per_row(),
                                             outer->threads(), (const uint64_t *)outer->Aj().elements(),
                                             outer->Ax().elements(), (const uint64_t *)outer->Arl().elements()),
                      "mg_set_outer");
        }
        info.device = h;
    }
    honei_hip_mg_config c{};
    c.n_max_iter = info.n_max_iter;
    c.n_pre_smooth = info.n_pre_smooth;
    c.n_post_smooth = info.n_post_smooth;
    c.initial_zero = info.initial_zero;
    c.convergence_check = info.convergence_check;
    c.prolmat = std::is_same<Prol_, methods::PROLMAT>::value ? 1 : 0;
    c.zero_inner_start = info.zero_inner_start;
    c.tolerance = (double)info.tolerance;
    c.adapt_correction_factor = (double)info.adapt_correction_factor;
    for (int i = 0; i < 8; ++i) c.mask[i] = info.macro_border_mask.elements()[i];
    c.max_outer = info.max_outer;
    c.post_smooth_uses_pre = 1;
    hip_check(honei_hip_mg_configure(info.device->h, &c), "mg_configure");
    return info.device->h;
}

}  // namespace hip

// Multigrid<..., JAC, V, MIXED>: FP32 inner V-cycles inside FP64 defect
// correction (math/multigrid.hh:2395-2536).
template <typename ProlType_>
struct Multigrid<tags::GPU::HIP, tags::GPU::HIP, ProlType_, methods::JAC, methods::CYCLE::V, methods::MIXED> {
    template <typename InnerPrec_, typename OuterPrec_>
    static void value(SparseMatrixELL<OuterPrec_> &system, DenseVector<OuterPrec_> &right_hand_side,
                      DenseVector<OuterPrec_> &x, unsigned long /*max_levels*/, OuterPrec_ /*conv_rad*/,
                      MGInfo<InnerPrec_, SparseMatrixELL<InnerPrec_>> &info) {
        static_assert(sizeof(OuterPrec_) == 8, "outer precision is double");
        hip::check_size(right_hand_side.size(), system.rows());
        hip::check_size(x.size(), system.rows());
        void *mg = hip::ensure_hierarchy<InnerPrec_, ProlType_>(info, &system);
        size_t outer = 0, inner = 0;
        double rd = 0;
        void *pb = hip::rd(right_hand_side);
        const int st = honei_hip_mg_solve_mixed(mg, (const double *)pb, (double *)hip::rw(x), &outer, &inner, &rd,
                                                nullptr);
        info.outer_iterations = outer;
        info.inner_iterations = inner;
        if (st != HONEI_HIP_ERROR_NOT_CONVERGED) hip_check(st, "Multigrid<tags::GPU::HIP, MIXED>");
    }
};

// Multigrid<..., JAC, V, FIXED> in double (math/multigrid.hh:828-870).
template <typename ProlType_>
struct Multigrid<tags::GPU::HIP, tags::GPU::HIP, ProlType_, methods::JAC, methods::CYCLE::V, methods::FIXED> {
    template <typename Prec_>
    static void value(SparseMatrixELL<Prec_> & /*system*/, DenseVector<Prec_> &right_hand_side,
                      DenseVector<Prec_> &x, unsigned long /*max_levels*/, Prec_ /*conv_rad*/,
                      MGInfo<Prec_, SparseMatrixELL<Prec_>> &info) {
        static_assert(sizeof(Prec_) == 8, "the device FIXED driver runs in double");
        void *mg = hip::ensure_hierarchy<Prec_, ProlType_>(info, nullptr);
        size_t iters = 0;
        double fd = 0;
        void *pb = hip::rd(right_hand_side);
        hip_check(honei_hip_mg_solve_fixed(mg, pb, hip::rw(x), &iters, &fd, nullptr), "Multigrid<tags::GPU::HIP, FIXED>");
        info.inner_iterations = iters;
    }
};

}  // namespace honei
