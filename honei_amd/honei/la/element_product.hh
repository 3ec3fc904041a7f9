// honei/la/element_product.hh -- ElementProduct<tags::GPU::HIP>
// (la/element_product.hh:611-647): r = a .* b.
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_> struct ElementProduct;

template <> struct ElementProduct<tags::GPU::HIP> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &r, const DenseVector<DT> &a, const DenseVector<DT> &b) {
        hip::check_size(a.size(), r.size());
        hip::check_size(b.size(), r.size());
        void *pa = hip::rd(a), *pb = hip::rd(b);
        void *pr = (r.same_storage(a) || r.same_storage(b)) ? hip::rw(r) : hip::wr(r);
        hip_check(honei_hip_element_product_three(hip::vt<DT>(), pr, pa, pb, r.size(), nullptr),
                  "ElementProduct<tags::GPU::HIP>");
        return r;
    }
};

}  // namespace honei
