// honei/math/jacobi.hh -- Jacobi<tags::GPU::HIP> multigrid overloads
// (math/jacobi.hh:481-513): the first sweep from zero is x = dinv .* rhs;
// further sweeps x <- x + dinv .* (rhs - A x), each one fused pass
// (math/jacobi.hh:268-279 runs Defect + ScaledSum).  diag_inverted already
// holds omega / diag(A) as the reference's callers build it.
#pragma once

#include "honei/backends/hip/device.hh"
#include "honei/la/element_product.hh"

namespace honei {

template <typename Tag_> struct Jacobi;

template <> struct Jacobi<tags::GPU::HIP> {
    template <typename DT>
    static void value(const SparseMatrixELL<DT> & /*A*/, const DenseVector<DT> &rhs, DenseVector<DT> &x,
                      DT /*omega*/, const DenseVector<DT> &diag_inverted) {
        ElementProduct<tags::GPU::HIP>::value(x, diag_inverted, rhs);
    }

    template <typename DT>
    static void value(DenseVector<DT> &to_smooth, const SparseMatrixELL<DT> &A, const DenseVector<DT> &rhs,
                      DenseVector<DT> &x, unsigned long iter_number, DT /*omega*/,
                      const DenseVector<DT> &diag_inverted) {
        if (iter_number == 0) {
            if (!x.same_storage(to_smooth)) copy_vec(x, to_smooth);
            return;
        }
        const DeviceEllMirror &m = A.device();
        DenseVector<DT> t(x.size());
        // ping-pong between to_smooth's data and t; result lands in x
        const DenseVector<DT> *src = &to_smooth;
        DenseVector<DT> *dst = (iter_number % 2 == 1) ? &x : &t;
        if (dst->same_storage(*src)) dst = &t;
        void *prhs = hip::rd(rhs), *pd = hip::rd(diag_inverted);
        for (unsigned long k = 0; k < iter_number; ++k) {
            void *ps = hip::rd(*src);
            void *pdst = hip::wr(*dst);
            hip_check(honei_hip_jacobi_smell_dv(m.vtype, pdst, ps, prhs, pd, m.Aj, m.Ax, m.Arl, A.rows(),
                                                A.num_cols_per_row(), A.stride(), A.threads(), nullptr, nullptr),
                      "Jacobi<tags::GPU::HIP>");
            src = dst;
            dst = (dst == &t) ? &x : &t;
            if (dst->same_storage(*src)) dst = &t;
        }
        if (!src->same_storage(x)) copy_vec(x, *src);
    }

  private:
    template <typename DT>
    static void copy_vec(DenseVector<DT> &dst, const DenseVector<DT> &src) {
        void *s = hip::rd(src);
        hip_check(honei_hip_copy(hip::wr(dst), s, sizeof(DT) * src.size(), nullptr), "Jacobi copy");
    }
};

// RISmoother<Tag> with a vector preconditioner (math/ri.hh:96-134):
// max_iters sweeps x += P .* (b - A x).
template <typename Tag_> struct RISmoother;

template <> struct RISmoother<tags::GPU::HIP> {
    static const unsigned long NUM_TEMPVECS = 2;
    template <typename DT>
    static void value(const SparseMatrixELL<DT> &A, const DenseVector<DT> &P, const DenseVector<DT> &b,
                      DenseVector<DT> &x, std::vector<DenseVector<DT>> & /*temp*/, unsigned long max_iters) {
        Jacobi<tags::GPU::HIP>::value(x, A, b, x, max_iters, DT(1), P);
    }
};

}  // namespace honei
