// honei/math/restriction.hh -- Restriction<tags::GPU::HIP, NONE | PROLMAT>
// (math/restriction.hh:40-200).  NONE: geometric full weighting on
// (2^L+1)^2 grids with the Dirichlet mask, bit-identical to the CPU sweep;
// PROLMAT: ELL product with resmat (= P^T).
#pragma once

#include "honei/backends/hip/device.hh"
#include "honei/la/product.hh"

namespace honei {

template <typename Tag_, typename Type_> struct Restriction;

template <> struct Restriction<tags::GPU::HIP, methods::NONE> {
    template <typename DT, typename MatrixType_>
    static DenseVector<DT> &value(DenseVector<DT> &coarse, const DenseVector<DT> &fine,
                                  const DenseVector<unsigned long> &mask, MatrixType_ & /*resmat*/) {
        const unsigned long *mk = mask.elements();
        void *pf = hip::rd(fine);
        hip_check(honei_hip_restriction(hip::vt<DT>(), hip::wr(coarse), coarse.size(), pf, fine.size(),
                                        (const uint64_t *)mk, nullptr),
                  "Restriction<tags::GPU::HIP, NONE>");
        return coarse;
    }
};

template <> struct Restriction<tags::GPU::HIP, methods::PROLMAT> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &coarse, const DenseVector<DT> &fine,
                                  const DenseVector<unsigned long> & /*mask*/, const SparseMatrixELL<DT> &resmat) {
        return Product<tags::GPU::HIP>::value(coarse, resmat, fine);
    }
};

}  // namespace honei
