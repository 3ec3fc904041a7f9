// honei/math/prolongation.hh -- Prolongation<tags::GPU::HIP, NONE | PROLMAT>
// (math/prolongation.hh:36-170): bilinear interpolation with the Dirichlet
// mask (bit-identical to the CPU sweep) or an ELL product with prolmat.
#pragma once

#include "honei/backends/hip/device.hh"
#include "honei/la/product.hh"

namespace honei {

template <typename Tag_, typename Type_> struct Prolongation;

template <> struct Prolongation<tags::GPU::HIP, methods::NONE> {
    template <typename DT, typename MatrixType_>
    static DenseVector<DT> &value(DenseVector<DT> &fine, const DenseVector<DT> &coarse,
                                  const DenseVector<unsigned long> &mask, MatrixType_ & /*prolmat*/) {
        const unsigned long *mk = mask.elements();
        void *pc = hip::rd(coarse);
        hip_check(honei_hip_prolongation(hip::vt<DT>(), hip::wr(fine), fine.size(), pc, coarse.size(),
                                         (const uint64_t *)mk, nullptr),
                  "Prolongation<tags::GPU::HIP, NONE>");
        return fine;
    }
};

template <> struct Prolongation<tags::GPU::HIP, methods::PROLMAT> {
    template <typename DT>
    static DenseVector<DT> &value(DenseVector<DT> &fine, const DenseVector<DT> &coarse,
                                  const DenseVector<unsigned long> & /*mask*/, const SparseMatrixELL<DT> &prolmat) {
        return Product<tags::GPU::HIP>::value(fine, prolmat, coarse);
    }
};

}  // namespace honei
