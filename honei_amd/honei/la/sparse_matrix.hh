// honei/la/sparse_matrix.hh -- row-wise sparse matrix used to build ELL
// matrices (the role of SparseMatrix<DT>, la/sparse_matrix.hh).  Rows keep
// their entries sorted by column; an assigned value is stored even when it
// is zero (SparseVector::used_elements counts it), as in the reference.
#pragma once

#include <map>
#include <vector>

namespace honei {

template <typename DT>
class SparseMatrix {
  public:
    SparseMatrix(unsigned long rows, unsigned long columns) : _rows(rows), _columns(columns), _r(rows) {}
    unsigned long rows() const { return _rows; }
    unsigned long columns() const { return _columns; }
    // (row, column, value): insert or overwrite
    void operator()(unsigned long row, unsigned long column, DT value) { _r[row][column] = value; }
    DT operator()(unsigned long row, unsigned long column) const {
        auto it = _r[row].find(column);
        return it == _r[row].end() ? DT(0) : it->second;
    }
    const std::map<unsigned long, DT> &row(unsigned long r) const { return _r[r]; }
    unsigned long used_elements() const {
        unsigned long n = 0;
        for (auto &m : _r) n += m.size();
        return n;
    }

  private:
    unsigned long _rows, _columns;
    std::vector<std::map<unsigned long, DT>> _r;
};

}  // namespace honei
