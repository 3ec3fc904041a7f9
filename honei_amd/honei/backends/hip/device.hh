// honei/backends/hip/device.hh -- helpers shared by the tags::GPU::HIP
// operator specialisations: device pointers of locked containers and the
// C-ABI type codes.
#pragma once

#include "honei/la/dense_vector.hh"
#include "honei/la/sparse_matrix_ell.hh"
#include "honei/util/configuration.hh"
#include "honei/util/exception.hh"
#include "honei_hip.h"

namespace honei {
namespace hip {

template <typename DT>
inline void *rd(const DenseVector<DT> &v) { return v.lock(lm_read_only, tags::tv_gpu_hip); }
template <typename DT>
inline void *wr(const DenseVector<DT> &v) { return v.lock(lm_write_only, tags::tv_gpu_hip); }
template <typename DT>
inline void *rw(const DenseVector<DT> &v) { return v.lock(lm_read_and_write, tags::tv_gpu_hip); }

template <typename DT> inline int vt() { return HipType<DT>::value; }

inline void check_size(unsigned long got, unsigned long want) {
    if (got != want) throw VectorSizeDoesNotMatch(got, want);
}

// Launch block size from the configuration (the reference reads
// cuda::product_smell_dv_* from honeirc); 0 = backend default.
inline unsigned long blocksize(const char *key) {
    return (unsigned long)Configuration::instance()->get_value(key, 0);
}

}  // namespace hip
}  // namespace honei
