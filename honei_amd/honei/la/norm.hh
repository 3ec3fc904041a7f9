// honei/la/norm.hh -- Norm<vnt_l_two, root, tags::GPU::HIP>
// (la/norm.hh:538-581): sum of squares (root = false) or its square root.
// The sum is formed on the device in double and returned to the host.
#pragma once

#include <cmath>

#include "honei/backends/hip/device.hh"

namespace honei {

template <VectorNormType norm_type_, bool root_, typename Tag_> struct Norm;

template <bool root_> struct Norm<vnt_l_two, root_, tags::GPU::HIP> {
    template <typename DT>
    static DT value(const DenseVector<DT> &x) {
        double r = 0.0;
        if (sizeof(DT) == 8) {
            hip_check(honei_hip_norm_l2_one_double(&r, hip::rd(x), x.size(), nullptr), "Norm<GPU::HIP>");
        } else {
            float f = 0.f;
            hip_check(honei_hip_norm_l2_one_float(&f, hip::rd(x), x.size(), nullptr), "Norm<GPU::HIP>");
            r = f;
        }
        return root_ ? DT(std::sqrt((DT)r)) : DT(r);
    }
};

}  // namespace honei
