// honei/math/transposition.hh -- Transposition of an ELL matrix through the
// row-wise SparseMatrix (math/transposition.hh:31-50); the result uses the
// configured ell::threads layout.
#pragma once

#include "honei/la/sparse_matrix_ell.hh"

namespace honei {

template <typename Tag_> struct Transposition {
    template <typename DT>
    static SparseMatrixELL<DT> value(const SparseMatrixELL<DT> &src) {
        SparseMatrix<DT> t(src.columns(), src.rows());
        const unsigned long *aj = src.Aj().elements();
        const DT *ax = src.Ax().elements();
        const unsigned long *arl = src.Arl().elements();
        const unsigned long T = src.threads(), S = src.stride();
        for (unsigned long r = 0; r < src.rows(); ++r)
            for (unsigned long k = 0; k < arl[r]; ++k)
                for (unsigned long l = 0; l < T; ++l) {
                    const unsigned long s = r * T + l + k * S;
                    if (ax[s] != DT(0)) t(aj[s], r, ax[s]);
                }
        return SparseMatrixELL<DT>(t);
    }
};

}  // namespace honei
