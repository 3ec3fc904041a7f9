// honei/math/cg.hh -- CGSolver<tags::GPU::HIP, methods::NONE>
// (math/cg.hh:63-133), ConjugateGradients<tags::GPU::HIP, NONE> with the
// konv_rad stopping rule (math/conjugate_gradients.hh:587-623) and
// ConjugateGradients<tags::GPU::HIP, JAC> (math/conjugate_gradients.hh:
// 1110-1154).
#pragma once

#include "honei/backends/hip/device.hh"

namespace honei {

template <typename Tag_, typename Method_> struct CGSolver;

template <> struct CGSolver<tags::GPU::HIP, methods::NONE> {
    // Whole solve on the device; scalars never leave it except for the
    // convergence flag read once per chunk of iterations.  A run that hits
    // max_iters returns normally with used_iters == max_iters, as the
    // reference does.
    template <typename DT, typename PreconContType_>
    static DenseVector<DT> &value(const SparseMatrixELL<DT> &A, const PreconContType_ & /*P*/,
                                  const DenseVector<DT> &b, DenseVector<DT> &x, unsigned long max_iters,
                                  unsigned long &used_iters, DT eps_relative = DT(1e-8)) {
        hip::check_size(b.size(), A.rows());
        hip::check_size(x.size(), A.rows());
        const DeviceEllMirror &m = A.device();
        size_t used = 0;
        double fd = 0;
        void *pb = hip::rd(b);
        const int st = honei_hip_cg_solve(m.vtype, m.Aj, m.Ax, m.Arl, A.rows(), A.num_cols_per_row(), A.stride(),
                                          A.threads(), pb, hip::rw(x), max_iters, (double)eps_relative, &used, &fd,
                                          nullptr);
        used_iters = used;
        if (st != HONEI_HIP_ERROR_NOT_CONVERGED) hip_check(st, "CGSolver<tags::GPU::HIP, NONE>");
        return x;
    }
};

template <typename Tag_, typename Method_> struct ConjugateGradients;

// Jacobi-preconditioned CG on ELL (math/conjugate_gradients.hh:1110-1154):
// the whole solve on the device, the convergence flag read once per chunk of
// iterations.  A run that hits max_iters returns normally with used_iters ==
// max_iters, as the reference does (which only prints then).
template <> struct ConjugateGradients<tags::GPU::HIP, methods::JAC> {
    template <typename DT>
    static void value(const SparseMatrixELL<DT> &A, const DenseVector<DT> &b, DenseVector<DT> &x,
                      const DenseVector<DT> &dd_inverted, unsigned long max_iters, unsigned long &used_iters,
                      DT eps_relative = DT(1e-8)) {
        hip::check_size(b.size(), A.rows());
        hip::check_size(x.size(), A.rows());
        hip::check_size(dd_inverted.size(), A.rows());
        const DeviceEllMirror &m = A.device();
        size_t used = 0;
        double fd = 0;
        void *pb = hip::rd(b);
        void *pd = hip::rd(dd_inverted);
        const int st = honei_hip_pcg_jacobi(m.vtype, m.Aj, m.Ax, m.Arl, A.rows(), A.num_cols_per_row(), A.stride(),
                                            A.threads(), pb, hip::rw(x), pd, max_iters, (double)eps_relative, &used,
                                            &fd, nullptr);
        if (max_iters > 0) used_iters = used;
        if (st != HONEI_HIP_ERROR_NOT_CONVERGED) hip_check(st, "ConjugateGradients<tags::GPU::HIP, JAC>");
    }
};

template <> struct ConjugateGradients<tags::GPU::HIP, methods::NONE> {
    // value(A, b, x, max_iters) (math/conjugate_gradients.hh:570-585, kernel
    // :204-249): CG from x with the fixed relative tolerance 1e-8 (and the
    // absolute stop ||r|| < 1e-8); the same iteration as CGSolver.
    template <typename DT>
    static void value(const SparseMatrixELL<DT> &A, const DenseVector<DT> &b, DenseVector<DT> &x,
                      unsigned long max_iters) {
        hip::check_size(b.size(), A.rows());
        hip::check_size(x.size(), A.rows());
        const DeviceEllMirror &m = A.device();
        size_t used = 0;
        double fd = 0;
        void *pb = hip::rd(b);
        const int st = honei_hip_cg_solve(m.vtype, m.Aj, m.Ax, m.Arl, A.rows(), A.num_cols_per_row(), A.stride(),
                                          A.threads(), pb, hip::rw(x), max_iters, 1e-8, &used, &fd, nullptr);
        if (st != HONEI_HIP_ERROR_NOT_CONVERGED) hip_check(st, "ConjugateGradients<tags::GPU::HIP, NONE>");
    }

    // value(A, b, x, konv_rad) (math/conjugate_gradients.hh:587-623), the
    // multigrid coarse solver: x = 0, iterate until ||x|| stagnates.
    template <typename DT>
    static void value(const SparseMatrixELL<DT> &A, const DenseVector<DT> &b, DenseVector<DT> &x, double konv_rad) {
        const DeviceEllMirror &m = A.device();
        size_t it = 0;
        void *pb = hip::rd(b);
        hip_check(honei_hip_cg_konv(m.vtype, m.Aj, m.Ax, m.Arl, A.rows(), A.num_cols_per_row(), A.stride(),
                                    A.threads(), pb, hip::wr(x), konv_rad, 4 * A.rows() + 64, &it, nullptr),
                  "ConjugateGradients<tags::GPU::HIP, NONE>");
    }
};

}  // namespace honei
