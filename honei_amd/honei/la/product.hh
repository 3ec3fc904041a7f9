// honei/la/product.hh -- Product<tags::GPU::HIP> (la/product.hh:1728-1794):
// y = A x for SparseMatrixELL, and the elementwise DenseVector overload.
#pragma once

#include "honei/backends/hip/device.hh"
#include "honei/la/element_product.hh"

namespace honei {

template <typename Tag_> struct Product;

template <> struct Product<tags::GPU::HIP> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &result, const SparseMatrixELL<DT> &a, const DenseVector<DT> &b) {
        hip::check_size(b.size(), a.columns());
        hip::check_size(result.size(), a.rows());
        const DeviceEllMirror &m = a.device();
        void *x = hip::rd(b);
        void *y = hip::wr(result);
        const unsigned long bs = hip::blocksize("hip::product_smell_dv");
        int st;
        if (m.window > 0)
            st = honei_hip_product_smell_dv_staged(m.vtype, x, y, m.Aj, m.Ax, m.Arl, a.rows(), a.num_cols_per_row(),
                                                   a.stride(), a.threads(), kStageBlock, 1, (const int32_t *)m.ranges,
                                                   m.nranges, m.window, nullptr);
        else if (m.vtype == HONEI_HIP_F64)
            st = honei_hip_product_smell_dv_double(x, y, m.Aj, m.Ax, m.Arl, 0, a.rows(), a.num_cols_per_row(),
                                                   a.stride(), bs, a.threads(), nullptr);
        else if (m.vtype == HONEI_HIP_BF16)
            st = honei_hip_product_smell_dv_bf16(x, y, m.Aj, m.Ax, m.Arl, 0, a.rows(), a.num_cols_per_row(),
                                                 a.stride(), bs, a.threads(), nullptr);
        else
            st = honei_hip_product_smell_dv_float(x, y, m.Aj, m.Ax, m.Arl, 0, a.rows(), a.num_cols_per_row(),
                                                  a.stride(), bs, a.threads(), nullptr);
        hip_check(st, "Product<tags::GPU::HIP>::value(ELL)");
        return result;
    }

    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &y, const DenseVector<DT> &a, const DenseVector<DT> &b) {
        return ElementProduct<tags::GPU::HIP>::value(y, a, b);
    }
};

}  // namespace honei
