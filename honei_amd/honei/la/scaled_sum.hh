// honei/la/scaled_sum.hh -- ScaledSum<tags::GPU::HIP> (la/scaled_sum.hh:444-487):
//   value(x, y, b)    x += b * y           (la/scaled_sum.hh:73-88)
//   value(r, x, y, b) r  = x + b * y       (la/scaled_sum.hh:91-110)
//   value(a, b, c)    a += b .* c          (la/scaled_sum.hh:177-196)
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_> struct ScaledSum;

template <> struct ScaledSum<tags::GPU::HIP> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &x, const DenseVector<DT> &y, DT b) {
        hip::check_size(y.size(), x.size());
        void *py = hip::rd(y);
        hip_check(honei_hip_scaled_sum_two(hip::vt<DT>(), hip::rw(x), py, (double)b, x.size(), nullptr),
                  "ScaledSum<GPU::HIP>(x, y, b)");
        return x;
    }
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &r, const DenseVector<DT> &x, const DenseVector<DT> &y, DT b) {
        hip::check_size(x.size(), r.size());
        hip::check_size(y.size(), r.size());
        void *px = hip::rd(x), *py = hip::rd(y);
        void *pr = (r.same_storage(x) || r.same_storage(y)) ? hip::rw(r) : hip::wr(r);
        hip_check(honei_hip_scaled_sum_three_s(hip::vt<DT>(), pr, px, py, (double)b, r.size(), nullptr),
                  "ScaledSum<GPU::HIP>(r, x, y, b)");
        return r;
    }
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &a, const DenseVector<DT> &b, const DenseVector<DT> &c) {
        hip::check_size(b.size(), a.size());
        hip::check_size(c.size(), a.size());
        void *pb = hip::rd(b), *pc = hip::rd(c);
        hip_check(honei_hip_scaled_sum_three(hip::vt<DT>(), hip::rw(a), pb, pc, a.size(), nullptr),
                  "ScaledSum<GPU::HIP>(a, b, c)");
        return a;
    }
};

}  // namespace honei
