// honei/la/element_inverse.hh -- ElementInverse<tags::GPU::HIP>: x = 1 / x.
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_> struct ElementInverse;

template <> struct ElementInverse<tags::GPU::HIP> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &x) {
        hip_check(honei_hip_element_inverse_one(hip::vt<DT>(), hip::rw(x), x.size(), nullptr),
                  "ElementInverse<tags::GPU::HIP>");
        return x;
    }
};

}  // namespace honei
