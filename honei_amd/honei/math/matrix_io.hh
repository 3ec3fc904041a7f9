// honei/math/matrix_io.hh -- MatrixIO<io_formats::ELL> binary format
// (math/matrix_io.hh:433-584): five uint64 header words (size, rows,
// columns, stride, num_cols_per_row), uint64 Aj[size], double Ax[size];
// T = 1 layout; row lengths derived from Ax.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "honei/la/sparse_matrix_ell.hh"

namespace honei {

template <typename IOFormat_> struct MatrixIO;

template <> struct MatrixIO<io_formats::ELL> {
    template <typename DT>
    static SparseMatrixELL<DT> read_matrix(const std::string &path, DT) {
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) throw InternalError("File " + path + " not found!");
        uint64_t h[5];
        if (std::fread(h, sizeof(uint64_t), 5, f) != 5) {
            std::fclose(f);
            throw InternalError("fread error!");
        }
        const uint64_t size = h[0];
        DenseVector<unsigned long> Aj(size);
        std::vector<double> ax(size);
        const bool ok = std::fread(Aj.elements(), sizeof(uint64_t), size, f) == size &&
                        std::fread(ax.data(), sizeof(double), size, f) == size;
        std::fclose(f);
        if (!ok) throw InternalError("fread error!");
        DenseVector<DT> Ax(size);
        DT *p = Ax.elements();
        for (uint64_t i = 0; i < size; ++i) p[i] = (DT)ax[i];
        return SparseMatrixELL<DT>(h[1], h[2], h[3], h[4], Aj, Ax, 1);
    }
};

}  // namespace honei
