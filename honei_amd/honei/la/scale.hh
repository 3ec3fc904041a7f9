// honei/la/scale.hh -- Scale<tags::GPU::HIP> (la/scale.hh:263-290): x *= a.
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_> struct Scale;

template <> struct Scale<tags::GPU::HIP> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &x, DT a) {
        hip_check(honei_hip_scale_one(hip::vt<DT>(), hip::rw(x), (double)a, x.size(), nullptr), "Scale<GPU::HIP>");
        return x;
    }
};

}  // namespace honei
