// honei/la/sum.hh -- Sum<tags::GPU::HIP> (la/sum.hh:684-728): x += y.
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_> struct Sum;

template <> struct Sum<tags::GPU::HIP> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &x, const DenseVector<DT> &y) {
        hip::check_size(y.size(), x.size());
        void *py = hip::rd(y);
        hip_check(honei_hip_sum_two(hip::vt<DT>(), hip::rw(x), py, x.size(), nullptr), "Sum<GPU::HIP>");
        return x;
    }
};

}  // namespace honei
