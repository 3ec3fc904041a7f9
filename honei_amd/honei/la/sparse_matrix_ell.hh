// honei/la/sparse_matrix_ell.hh -- SparseMatrixELL<DT> (la/sparse_matrix_ell.hh:44-129)
// with a device mirror for tags::GPU::HIP.
//
// Host layout and construction rules are the reference's
// (la/sparse_matrix_ell-impl.hh:75-208): T = `threads` lanes per row, entry
// t of row r at slot (t % T) + r*T + (t / T)*stride, stride 32-aligned,
// num_cols_per_row = ceil(max row length / T) >= 1, Arl[r] = ceil(len_r / T),
// zero values dropped, zero padding.  The device mirror holds 32-bit column
// indices and row lengths (narrowed and range-checked on upload) and the
// values in DT, or in bf16 when requested for FP32 matrices.
#pragma once

#include <cmath>
#include <memory>
#include <vector>

#include "honei/la/dense_vector.hh"
#include "honei/la/sparse_matrix.hh"
#include "honei/util/configuration.hh"
#include "honei_hip.h"

namespace honei {

template <typename DT> struct HipType;
template <> struct HipType<float> { static const int value = HONEI_HIP_F32; };
template <> struct HipType<double> { static const int value = HONEI_HIP_F64; };

// Device-resident copy of an ELL matrix (what the kernels read).
struct DeviceEllMirror {
    void *Aj = nullptr, *Ax = nullptr, *Arl = nullptr;
    int vtype = HONEI_HIP_F32;
    // x staged in LDS (honei_hip_product_smell_dv_staged): column windows of
    // every stage_rows rows, their count and the LDS elements to reserve
    // (0: staging not used)
    void *ranges = nullptr;
    unsigned long nranges = 0;
    long long window = 0;
    ~DeviceEllMirror() {
        if (Aj) honei_hip_free(Aj);
        if (Ax) honei_hip_free(Ax);
        if (Arl) honei_hip_free(Arl);
        if (ranges) honei_hip_free(ranges);
    }
};

// Work-group shape of the staged product: 1024 threads, one chunk.
constexpr unsigned long kStageBlock = 1024;

template <typename DT>
class SparseMatrixELL {
    struct Imp {
        unsigned long rows, columns, stride, num_cols_per_row, threads;
        DenseVector<unsigned long> Aj, Arl;
        DenseVector<DT> Ax;
        std::shared_ptr<DeviceEllMirror> dev;
        bool bf16 = false, stage_x = false;
        Imp(unsigned long r, unsigned long c, unsigned long s, unsigned long n, unsigned long t,
            const DenseVector<unsigned long> &aj, const DenseVector<DT> &ax, const DenseVector<unsigned long> &arl)
            : rows(r), columns(c), stride(s), num_cols_per_row(n), threads(t), Aj(aj), Arl(arl), Ax(ax) {}
    };
    std::shared_ptr<Imp> _imp;

    static unsigned long ceil_div(unsigned long a, unsigned long b) {
        return (unsigned long)std::ceil(double(a) / double(b));
    }

  public:
    typedef DT iDT_;

    // Explicit arrays; row lengths derived from Ax
    // (la/sparse_matrix_ell-impl.hh:61-73, 186-208).
    SparseMatrixELL(unsigned long rows, unsigned long columns, unsigned long stride, unsigned long num_cols_per_row,
                    const DenseVector<unsigned long> &Aj, const DenseVector<DT> &Ax, unsigned long threads)
        : _imp(std::make_shared<Imp>(rows, columns, stride, num_cols_per_row, threads, Aj, Ax,
                                     DenseVector<unsigned long>(rows, 0ul))) {
        const DT *ax = Ax.elements();
        unsigned long *arl = _imp->Arl.elements();
        for (unsigned long row = 0; row < rows; ++row) {
            unsigned long count = 0;
            bool stop = false;
            for (unsigned long i = row * threads; i < Ax.size() && !stop; i += stride)
                for (unsigned long t = 0; t < threads; ++t) {
                    if (ax[i + t] == DT(0)) {
                        stop = true;
                        break;
                    }
                    if (t == 0) ++count;
                }
            arl[row] = count;
        }
    }

    // From a SparseMatrix (la/sparse_matrix_ell-impl.hh:75-122), T from the
    // configuration key ell::threads.
    explicit SparseMatrixELL(const SparseMatrix<DT> &src)
        : _imp(std::make_shared<Imp>(src.rows(), src.columns(), 0, 1,
                                     (unsigned long)Configuration::instance()->get_value("ell::threads", 1),
                                     DenseVector<unsigned long>(1), DenseVector<DT>(1),
                                     DenseVector<unsigned long>(src.rows(), 0ul))) {
        const unsigned long T = _imp->threads, rows = src.rows();
        unsigned long width = 1;
        unsigned long *arl = _imp->Arl.elements();
        for (unsigned long r = 0; r < rows; ++r) {
            const unsigned long ue = src.row(r).size();
            arl[r] = ceil_div(ue, T);
            if (ue > width) width = ue;
        }
        _imp->num_cols_per_row = ceil_div(width, T);
        _imp->stride = 32 * ((rows * T + 31) / 32);
        const unsigned long size = _imp->num_cols_per_row * _imp->stride;
        DenseVector<unsigned long> Aj(size, 0ul);
        DenseVector<DT> Ax(size, DT(0));
        unsigned long *aj = Aj.elements();
        DT *ax = Ax.elements();
        for (unsigned long r = 0; r < rows; ++r) {
            unsigned long target = 0;
            for (auto &kv : src.row(r)) {
                if (kv.second == DT(0)) continue;
                const unsigned long s = (target % T) + r * T + (target / T) * _imp->stride;
                aj[s] = kv.first;
                ax[s] = kv.second;
                ++target;
            }
        }
        _imp->Aj = Aj;
        _imp->Ax = Ax;
    }

    SparseMatrixELL(const SparseMatrixELL &other) = default;
    SparseMatrixELL &operator=(const SparseMatrixELL &other) = default;

    unsigned long rows() const { return _imp->rows; }
    unsigned long columns() const { return _imp->columns; }
    unsigned long size() const { return _imp->rows * _imp->columns; }
    unsigned long stride() const { return _imp->stride; }
    unsigned long num_cols_per_row() const { return _imp->num_cols_per_row; }
    unsigned long threads() const { return _imp->threads; }
    DenseVector<unsigned long> &Aj() const { return _imp->Aj; }
    DenseVector<DT> &Ax() const { return _imp->Ax; }
    DenseVector<unsigned long> &Arl() const { return _imp->Arl; }
    unsigned long used_elements() const {
        unsigned long ue = 0;
        const DT *ax = _imp->Ax.elements();
        for (unsigned long i = 0; i < _imp->Ax.size(); ++i)
            if (ax[i] != DT(0)) ++ue;
        return ue;
    }
    // Element (row, column) (la/sparse_matrix_ell-impl.hh:326-338).
    DT operator()(unsigned long row, unsigned long column) const {
        const unsigned long *aj = _imp->Aj.elements();
        const DT *ax = _imp->Ax.elements();
        const unsigned long max = _imp->Arl.elements()[row];
        for (unsigned long i = row * threads(), j = 0; j < max && aj[i] <= column; i += stride(), ++j)
            for (unsigned long t = 0; t < threads(); ++t)
                if (aj[i + t] == column) return ax[i + t];
        return DT(0);
    }

    // Store the device copy of the values as bf16 (FP32 matrices only);
    // arithmetic stays FP32.
    void set_device_bf16(bool on) {
        _imp->bf16 = on;
        _imp->dev.reset();
    }
    // Products gather x from an LDS copy of each work-group's column window
    // (pays off for scattered columns inside a band; same results).  The
    // windows are computed with the device copy; a matrix whose widest
    // window does not fit keeps the plain product.
    void set_stage_x(bool on) {
        _imp->stage_x = on;
        _imp->dev.reset();
    }
    bool stage_x() const { return _imp->stage_x; }
    // Drop the device copy (after host-side modification of Aj/Ax/Arl).
    void invalidate_device() const { _imp->dev.reset(); }

    // Device mirror, uploaded on first use.
    const DeviceEllMirror &device() const {
        if (!_imp->dev) {
            auto m = std::make_shared<DeviceEllMirror>();
            const unsigned long size = _imp->Ax.size();
            hip_check(honei_hip_malloc(&m->Aj, 4 * (size ? size : 1)), "ELL Aj allocation");
            hip_check(honei_hip_malloc(&m->Arl, 4 * (rows() ? rows() : 1)), "ELL Arl allocation");
            hip_check(honei_hip_upload_indices(m->Aj, (const uint64_t *)_imp->Aj.elements(), size, columns()),
                      "ELL Aj upload");
            hip_check(honei_hip_upload_indices(m->Arl, (const uint64_t *)_imp->Arl.elements(), rows(),
                                               num_cols_per_row() + 1),
                      "ELL Arl upload");
            if (_imp->bf16 && sizeof(DT) == 4) {
                m->vtype = HONEI_HIP_BF16;
                hip_check(honei_hip_malloc(&m->Ax, 2 * (size ? size : 1)), "ELL Ax allocation");
                hip_check(honei_hip_upload_values_bf16(m->Ax, (const float *)_imp->Ax.elements(), size),
                          "ELL Ax upload");
            } else {
                m->vtype = HipType<DT>::value;
                hip_check(honei_hip_malloc(&m->Ax, sizeof(DT) * (size ? size : 1)), "ELL Ax allocation");
                hip_check(honei_hip_upload(m->Ax, _imp->Ax.elements(), sizeof(DT) * size, nullptr), "ELL Ax upload");
            }
            if (_imp->stage_x && rows() > 0) {
                const unsigned long rpb = kStageBlock / threads();
                m->nranges = (rows() + rpb - 1) / rpb;
                hip_check(honei_hip_malloc(&m->ranges, 8 * m->nranges), "ELL column window allocation");
                int64_t w = 0;
                hip_check(honei_hip_ell_block_ranges(m->Aj, m->Arl, rows(), num_cols_per_row(), stride(), threads(), rpb,
                                                     (int32_t *)m->ranges, &w, nullptr),
                          "ELL column windows");
                const long long cap = (96 * 1024) / (m->vtype == HONEI_HIP_F64 ? 8 : 4);
                m->window = (w > 0 && w <= cap) ? w : 0;
            }
            _imp->dev = m;
        }
        return *_imp->dev;
    }

    SparseMatrixELL copy() const {
        return SparseMatrixELL(rows(), columns(), stride(), num_cols_per_row(), _imp->Aj.copy(), _imp->Ax.copy(),
                               threads());
    }
};

template <typename DT>
bool operator==(const SparseMatrixELL<DT> &a, const SparseMatrixELL<DT> &b) {
    if (a.columns() != b.columns() || a.rows() != b.rows()) return false;
    if (a.Ax().size() != b.Ax().size()) return false;
    const unsigned long *aj = a.Aj().elements(), *bj = b.Aj().elements();
    const DT *ax = a.Ax().elements(), *bx = b.Ax().elements();
    for (unsigned long i = 0; i < a.Ax().size(); ++i)
        if (aj[i] != bj[i] || ax[i] != bx[i]) return false;
    return true;
}

}  // namespace honei
