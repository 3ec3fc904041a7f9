// honei/la/algorithm.hh -- copy / fill / convert on tags::GPU::HIP
// (la/algorithm.hh:50-198).
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_, typename DT>
void copy(const DenseVector<DT> &source, DenseVector<DT> &dest) {
    hip::check_size(dest.size(), source.size());
    void *s = hip::rd(source);
    hip_check(honei_hip_copy(hip::wr(dest), s, sizeof(DT) * source.size(), nullptr), "copy<GPU::HIP>");
}

template <typename Tag_, typename DT>
void fill(DenseVector<DT> &x, DT value = DT(0)) {
    if (value == DT(0)) {
        hip_check(honei_hip_fill_zero(hip::wr(x), sizeof(DT) * x.size(), nullptr), "fill<GPU::HIP>");
    } else {
        hip_check(honei_hip_fill(hip::vt<DT>(), hip::wr(x), (double)value, x.size(), nullptr), "fill<GPU::HIP>");
    }
}

// convert<Tag>(dest, src): double <-> float (la/algorithm.hh:175-198)
template <typename Tag_>
void convert(DenseVector<float> &dest, const DenseVector<double> &src) {
    hip::check_size(dest.size(), src.size());
    void *s = hip::rd(src);
    hip_check(honei_hip_convert_double_float((float *)hip::wr(dest), (const double *)s, src.size(), nullptr),
              "convert<GPU::HIP>(float, double)");
}
template <typename Tag_>
void convert(DenseVector<double> &dest, const DenseVector<float> &src) {
    hip::check_size(dest.size(), src.size());
    void *s = hip::rd(src);
    hip_check(honei_hip_convert_float_double((double *)hip::wr(dest), (const float *)s, src.size(), nullptr),
              "convert<GPU::HIP>(double, float)");
}

}  // namespace honei
