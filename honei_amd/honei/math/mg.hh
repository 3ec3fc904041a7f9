// honei/math/mg.hh -- operator-list multigrid (math/mg.hh:35-1127) running
// its operators on tags::GPU::HIP.
//
// MGData holds per-level systems, transfers, Jacobi preconditioners and
// vectors (internal index = level - 1 for A/vectors, level - 2 for
// transfers, MGDataIndex of the reference).  MGCycleCreation builds a flat
// list of operations for one V, W or F cycle; MGSolver applies it until the
// relative defect drops below eps.  Every operation is a device operator of
// this backend, so vectors stay in HBM for the whole solve.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "honei/la/algorithm.hh"
#include "honei/la/norm.hh"
#include "honei/la/product.hh"
#include "honei/la/sum.hh"
#include "honei/math/cg.hh"
#include "honei/math/defect.hh"
#include "honei/math/jacobi.hh"
#include "honei/math/matrix_io.hh"
#include "honei/math/prolongation.hh"
#include "honei/math/restriction.hh"
#include "honei/math/transposition.hh"
#include "honei/math/vector_io.hh"

namespace honei {

struct MGDataIndex {
    static unsigned long internal_index_A(unsigned long base) { return base - 1; }
    static unsigned long internal_index_Prol(unsigned long base) { return base - 2; }
};

template <typename MatrixType_, typename VectorType_, typename TransferContType_, typename PreconContType_,
          typename DataType_>
struct MGData {
    std::vector<MatrixType_> A;
    std::vector<TransferContType_> resmat, prolmat;
    std::vector<PreconContType_> P;
    std::vector<VectorType_> b, x, d, c, store;
    std::vector<std::vector<VectorType_>> smoother_temp;
    unsigned long max_iters = 0, max_iters_coarse = 0, used_iters_coarse = 0, used_iters = 0;
    unsigned long n_pre_smooth = 0, n_post_smooth = 0, min_level = 1;
    DataType_ eps_relative = DataType_(0);
};

// A flat list of operations for one cycle (math/operator_list.hh:31-88).
class OperatorList {
  public:
    void push_back(std::function<void()> op, char kind) {
        _ops.push_back(std::move(op));
        _kinds.push_back(kind);
    }
    void value() {
        for (auto &op : _ops) op();
    }
    unsigned long size() const { return _ops.size(); }
    // S smoother, D defect, R restriction, F fill, P prolongation, + sum, C coarse solve
    std::string shape() const { return _kinds; }

  private:
    std::vector<std::function<void()>> _ops;
    std::string _kinds;
};

template <typename Tag_, typename MatrixType_, typename VectorType_, typename TransferContType_,
          typename PreconContType_, typename MatIOType_, typename VecIOType_, typename DataType_>
struct MGUtil {
    typedef MGData<MatrixType_, VectorType_, TransferContType_, PreconContType_, DataType_> Data;

    static void configure(Data &t, unsigned long max_iters, unsigned long max_iters_coarse,
                          unsigned long n_pre_smooth, unsigned long n_post_smooth, unsigned long min_level,
                          DataType_ eps_relative) {
        t.max_iters = max_iters;
        t.max_iters_coarse = max_iters_coarse;
        t.n_pre_smooth = n_pre_smooth;
        t.n_post_smooth = n_post_smooth;
        t.min_level = min_level;
        t.eps_relative = eps_relative;
    }

    // Files A_<l>.ell (l = 1..max_level), prol_<l>.ell (l = 2..max_level),
    // rhs_<max>, init_<max> (math/mg.hh:442-545).  "jac" preconditioner:
    // damping / diag(A).  Restriction matrices are the transposed
    // prolongations.
    static Data load_data(const std::string &base, unsigned long max_level, DataType_ damping,
                          const std::string &precon) {
        if (precon != "jac") throw InternalError("Preconditioner unknown!");
        Data d;
        for (unsigned long i = 0; i < max_level; ++i) {
            MatrixType_ A(MatIOType_::read_matrix(base + "A_" + std::to_string(i + 1) + ".ell", DataType_(0)));
            PreconContType_ P(A.rows(), damping);
            DataType_ *p = P.elements();
            for (unsigned long j = 0; j < P.size(); ++j) p[j] /= A(j, j);
            d.A.push_back(A);
            d.P.push_back(P);
            if (i + 1 < max_level) {
                TransferContType_ pr(
                    MatIOType_::read_matrix(base + "prol_" + std::to_string(i + 2) + ".ell", DataType_(0)));
                d.prolmat.push_back(pr);
                d.resmat.push_back(Transposition<Tag_>::value(pr));
            }
            if (i + 1 == max_level) {
                d.b.push_back(VecIOType_::read_vector(base + "rhs_" + std::to_string(max_level), DataType_(0)));
                d.x.push_back(VecIOType_::read_vector(base + "init_" + std::to_string(max_level), DataType_(0)));
            } else {
                d.b.push_back(VectorType_(A.rows(), DataType_(0)));
                d.x.push_back(VectorType_(A.rows(), DataType_(0)));
            }
            d.d.push_back(VectorType_(A.rows(), DataType_(0)));
            d.c.push_back(VectorType_(A.rows(), DataType_(0)));
            d.store.push_back(VectorType_(A.rows(), DataType_(0)));
            d.smoother_temp.push_back(std::vector<VectorType_>());
        }
        return d;
    }
};

template <typename Tag_, typename CycleShape_, typename CoarseGridSolverType_, typename SmootherType_,
          typename ResType_, typename ProlType_, typename DataType_>
struct MGCycleCreation {
    template <typename M, typename V, typename T, typename P>
    static OperatorList value(MGData<M, V, T, P, DataType_> &data) {
        OperatorList list;
        build(list, data, data.x, data.b, data.A.size(), shape());
        return list;
    }

  private:
    enum Shape { kV, kW, kF };
    static Shape shape() {
        if (std::is_same<CycleShape_, methods::CYCLE::W::STATIC>::value) return kW;
        if (std::is_same<CycleShape_, methods::CYCLE::F::V::STATIC>::value) return kF;
        return kV;
    }

    template <typename Data, typename Vecs>
    static void build(OperatorList &list, Data &data, Vecs &x, Vecs &b, unsigned long level, Shape s) {
        const unsigned long ia = MGDataIndex::internal_index_A(level);
        if (level == data.min_level) {
            auto A = data.A.at(ia);
            auto P = data.P.at(ia);
            auto bb = b.at(ia);
            auto xx = x.at(ia);
            Data *dp = &data;
            list.push_back([A, P, bb, xx, dp]() mutable {
                CoarseGridSolverType_::value(A, P, bb, xx, dp->max_iters_coarse, dp->used_iters_coarse,
                                             dp->eps_relative);
            }, 'C');
            return;
        }
        auto A = data.A.at(ia);
        auto P = data.P.at(ia);
        auto bb = b.at(ia);
        auto xx = x.at(ia);
        auto st = data.store.at(ia);
        auto tmp = data.smoother_temp.at(ia);
        auto dc = data.d.at(ia - 1);
        auto cc = data.c.at(ia - 1);
        auto R = data.resmat.at(MGDataIndex::internal_index_Prol(level));
        auto Pr = data.prolmat.at(MGDataIndex::internal_index_Prol(level));
        const unsigned long npre = data.n_pre_smooth, npost = data.n_post_smooth;
        DenseVector<unsigned long> mask(8, 2ul);
        list.push_back([A, P, bb, xx, tmp, npre]() mutable { SmootherType_::value(A, P, bb, xx, tmp, npre); }, 'S');
        list.push_back([st, bb, A, xx]() mutable { Defect<Tag_>::value(st, bb, A, xx); }, 'D');
        list.push_back([dc, st, mask, R]() mutable { ResType_::value(dc, st, mask, R); }, 'R');
        list.push_back([cc]() mutable { fill<Tag_>(cc, DataType_(0)); }, 'F');
        if (s == kV) {
            build(list, data, data.c, data.d, level - 1, kV);
        } else if (s == kW) {
            if (level - 1 != data.min_level) build(list, data, data.c, data.d, level - 1, kW);
            build(list, data, data.c, data.d, level - 1, kW);
        } else {
            build(list, data, data.c, data.d, level - 1, kF);
            build(list, data, data.c, data.d, level - 1, kV);
        }
        list.push_back([st, cc, mask, Pr]() mutable { ProlType_::value(st, cc, mask, Pr); }, 'P');
        list.push_back([xx, st]() mutable { Sum<Tag_>::value(xx, st); }, '+');
        list.push_back([A, P, bb, xx, tmp, npost]() mutable { SmootherType_::value(A, P, bb, xx, tmp, npost); }, 'S');
    }
};

// MGSolver<Tag, Norm> (math/mg.hh:1093-1127): cycles until the defect norm
// drops below eps_relative times the initial one.
template <typename Tag_, typename NormType_>
struct MGSolver {
    template <typename M, typename V, typename T, typename P, typename DT>
    static void value(MGData<M, V, T, P, DT> &data, OperatorList &cycle) {
        if (cycle.size() == 0) throw InternalError("OperatorList is empty!");
        V &b = data.b.back();
        V &x = data.x.back();
        M &A = data.A.back();
        V r(x.size());
        Defect<Tag_>::value(r, b, A, x);
        const DT initial = NormType_::value(r);
        for (unsigned long i = 0; i < data.max_iters; ++i) {
            cycle.value();
            Defect<Tag_>::value(r, b, A, x);
            const DT current = NormType_::value(r);
            data.used_iters = i + 1;
            if (current < data.eps_relative * initial) break;
        }
    }
};

}  // namespace honei
