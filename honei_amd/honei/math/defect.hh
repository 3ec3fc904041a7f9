// honei/math/defect.hh -- Defect<tags::GPU::HIP> (math/defect.hh:645-692):
// r = rhs - A x on ELL, one fused pass.
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_> struct Defect;

template <> struct Defect<tags::GPU::HIP> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &r, const DenseVector<DT> &rhs, const SparseMatrixELL<DT> &A,
                                  const DenseVector<DT> &x) {
        hip::check_size(x.size(), A.columns());
        hip::check_size(rhs.size(), A.columns());
        hip::check_size(r.size(), A.rows());
        const DeviceEllMirror &m = A.device();
        void *px = hip::rd(x), *pb = hip::rd(rhs);
        void *pr = (r.same_storage(x) || r.same_storage(rhs)) ? nullptr : hip::wr(r);
        if (!pr) {   // in-place request: go through a temporary (rows read x of other rows)
            DenseVector<DT> t(r.size());
            value(t, rhs, A, x);
            copy_into(r, t);
            return r;
        }
        const unsigned long bs = hip::blocksize("hip::defect_smell_dv");
        int st;
        if (m.vtype == HONEI_HIP_F64)
            st = honei_hip_defect_smell_dv_double(pb, pr, m.Aj, m.Ax, m.Arl, px, 0, A.rows(), A.num_cols_per_row(),
                                                  A.stride(), bs, A.threads(), nullptr);
        else if (m.vtype == HONEI_HIP_F32)
            st = honei_hip_defect_smell_dv_float(pb, pr, m.Aj, m.Ax, m.Arl, px, 0, A.rows(), A.num_cols_per_row(),
                                                 A.stride(), bs, A.threads(), nullptr);
        else   // bf16 values: fused defect kernel with FP32 vectors
            st = honei_hip_defect_norm_smell_dv(HONEI_HIP_BF16, pb, pr, m.Aj, m.Ax, m.Arl, px, A.rows(),
                                                A.num_cols_per_row(), A.stride(), A.threads(), scratch_norm(), nullptr);
        hip_check(st, "Defect<tags::GPU::HIP>");
        return r;
    }

    template <typename DT>
    static DenseVector<DT> value(const DenseVector<DT> &rhs, const SparseMatrixELL<DT> &A, const DenseVector<DT> &x) {
        DenseVector<DT> r(rhs.size());
        value(r, rhs, A, x);
        return r;
    }

  private:
    template <typename DT>
    static void copy_into(DenseVector<DT> &dst, const DenseVector<DT> &src) {
        void *s = hip::rd(src);
        hip_check(honei_hip_copy(hip::wr(dst), s, sizeof(DT) * src.size(), nullptr), "Defect copy");
    }
    static double *scratch_norm() {
        static DenseVector<double> n(1);
        return (double *)n.lock(lm_write_only, tags::tv_gpu_hip);
    }
};

}  // namespace honei
