// honei/la/difference.hh -- Difference<tags::GPU::HIP> (la/difference.hh:710-728):
// value(x, y): x -= y ; value(r, x, y): r = x - y.
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_> struct Difference;

template <> struct Difference<tags::GPU::HIP> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &x, const DenseVector<DT> &y) {
        hip::check_size(y.size(), x.size());
        void *py = hip::rd(y);
        hip_check(honei_hip_difference_two(hip::vt<DT>(), hip::rw(x), py, x.size(), nullptr), "Difference<GPU::HIP>");
        return x;
    }
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &r, const DenseVector<DT> &x, const DenseVector<DT> &y) {
        hip::check_size(x.size(), r.size());
        hip::check_size(y.size(), r.size());
        void *px = hip::rd(x), *py = hip::rd(y);
        void *pr = (r.same_storage(x) || r.same_storage(y)) ? hip::rw(r) : hip::wr(r);
        hip_check(honei_hip_difference_three(hip::vt<DT>(), pr, px, py, r.size(), nullptr), "Difference<GPU::HIP>");
        return r;
    }
};

}  // namespace honei
