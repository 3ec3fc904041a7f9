// convert_tool -- runs the C++ mirror's SparseMatrix -> SparseMatrixELL
// conversion (honei_amd/honei/la/sparse_matrix_ell.hh; the reference's
// la/sparse_matrix_ell-impl.hh:75-122) on the host, for the CPU test that
// compares its index and value arrays bit for bit with the reference's own
// files and build (tests/test_cpp_convert.py).  No device work.
//
// stdin : uint64 rows, columns, threads, entries; then per entry uint64 row,
//         uint64 column, double value (assigned in this order).
// stdout: uint64 stride, num_cols_per_row; Aj (uint64 [ncpr*stride]);
//         Ax (double [ncpr*stride]); Arl (uint64 [rows]).
#include <cstdint>
#include <cstdio>
#include <vector>

#include "honei/la/sparse_matrix_ell.hh"

static bool rd(void *p, size_t bytes) { return std::fread(p, 1, bytes, stdin) == bytes; }
static void wr(const void *p, size_t bytes) { std::fwrite(p, 1, bytes, stdout); }

int main() {
    uint64_t h[4];
    if (!rd(h, sizeof(h))) return 2;
    const uint64_t rows = h[0], cols = h[1], threads = h[2], entries = h[3];
    honei::SparseMatrix<double> sm(rows, cols);
    for (uint64_t e = 0; e < entries; ++e) {
        uint64_t rc[2];
        double v;
        if (!rd(rc, sizeof(rc)) || !rd(&v, sizeof(v))) return 3;
        sm(rc[0], rc[1], v);
    }
    honei::Configuration::instance()->set_value("ell::threads", (int)threads);
    honei::SparseMatrixELL<double> ell(sm);
    const uint64_t hdr[2] = {ell.stride(), ell.num_cols_per_row()};
    wr(hdr, sizeof(hdr));
    const unsigned long n = ell.Ax().size();
    wr(ell.Aj().elements(), 8 * n);
    wr(ell.Ax().elements(), 8 * n);
    wr(ell.Arl().elements(), 8 * rows);
    return std::fflush(stdout) == 0 ? 0 : 4;
}
