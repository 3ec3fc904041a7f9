// honei/la/dot_product.hh -- DotProduct<tags::GPU::HIP> (la/dot_product.hh:343-368).
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_> struct DotProduct;

template <> struct DotProduct<tags::GPU::HIP> {
    template <typename DT>
    static DT value(const DenseVector<DT> &x, const DenseVector<DT> &y) {
        hip::check_size(y.size(), x.size());
        if (sizeof(DT) == 8) {
            double r = 0;
            hip_check(honei_hip_dot_product_two_double(&r, hip::rd(x), hip::rd(y), x.size(), nullptr),
                      "DotProduct<GPU::HIP>");
            return DT(r);
        }
        float f = 0;
        hip_check(honei_hip_dot_product_two_float(&f, hip::rd(x), hip::rd(y), x.size(), nullptr),
                  "DotProduct<GPU::HIP>");
        return DT(f);
    }
};

}  // namespace honei
