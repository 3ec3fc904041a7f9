// ref_driver.cc -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" harness around HONEI's own tags::CPU code, compiled (by
// `make -C oracle ref`) together with the reference's sources where they lie
// under /root/reference.  The output library, oracle/_ref/libhonei_ref.so,
// is the reference itself; tests use it to pin the clean-room restatement in
// honei_oracle.cpp and bench.py may time it as the CPU baseline.  No
// reference source is copied here: this file only calls the reference's API.

#include <honei/la/banded_matrix_qx.hh>
#include <honei/la/dense_vector.hh>
#include <honei/la/product.hh>
#include <honei/la/sparse_matrix.hh>
#include <honei/la/sparse_matrix_csr.hh>
#include <honei/la/sparse_matrix_ell.hh>
#include <honei/math/cg.hh>
#include <honei/math/conjugate_gradients.hh>
#include <honei/math/defect.hh>
#include <honei/math/fill_matrix.hh>
#include <honei/math/multigrid.hh>
#include <honei/math/prolongation.hh>
#include <honei/math/prolongation_matrix.hh>
#include <honei/math/restriction.hh>
#include <honei/util/configuration.hh>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

using namespace honei;
typedef uint64_t u64;

namespace {

struct HandleF { SparseMatrixELL<float> *m; };
struct HandleD { SparseMatrixELL<double> *m; };

template <typename DT> struct H;
template <> struct H<float> { typedef HandleF type; };
template <> struct H<double> { typedef HandleD type; };

void set_threads(u64 threads) {
    Configuration::instance()->set_value("ell::threads", (int)threads);
}

template <typename DT>
SparseMatrix<DT> sm_from_csr(u64 rows, u64 cols, const u64 *rp, const u64 *col, const DT *val) {
    SparseMatrix<DT> sm(rows, cols);
    for (u64 r = 0; r < rows; ++r)
        for (u64 k = rp[r]; k < rp[r + 1]; ++k) sm(r, col[k], val[k]);
    return sm;
}

template <typename DT>
SparseMatrixELL<DT> *q1_ell(u64 side, int bc) {
    const u64 n = side * side;
    DenseVector<DT> band(n, DT(0));
    BandedMatrixQx<Q1Type, DT> bm(n, band.copy(), band.copy(), band.copy(), band.copy(), band.copy(),
                                  band.copy(), band.copy(), band.copy(), band.copy());
    if (bc == 0)
        FillMatrix<tags::CPU, applications::POISSON, boundary_types::DIRICHLET::DIRICHLET_0>::value(bm);
    else
        FillMatrix<tags::CPU, applications::POISSON, boundary_types::DIRICHLET_NEUMANN>::value(bm);
    SparseMatrix<DT> sm(bm);
    return new SparseMatrixELL<DT>(sm);
}

template <typename DT>
void copy_out(const DenseVector<DT> &v, DT *dst) {
    for (u64 i = 0; i < v.size(); ++i) dst[i] = v[i];
}

// Capture std::cout (the MG drivers print their iteration counts).
struct CoutCapture {
    std::stringstream ss;
    std::streambuf *old;
    CoutCapture() : old(std::cout.rdbuf(ss.rdbuf())) {}
    ~CoutCapture() { std::cout.rdbuf(old); }
    u64 find_count(const std::string &key) const {
        const std::string s = ss.str();
        size_t p = s.rfind(key);
        if (p == std::string::npos) return 0;
        return (u64)std::strtoull(s.c_str() + p + key.size(), nullptr, 10);
    }
};

}  // namespace

// ---- banded Q1: Product<tags::CPU>::value(BandedMatrixQx, DenseVector) and
// Defect<tags::CPU> on it; bands[k*n + i], k = LL, LD, LU, DL, DD, DU, UL, UD, UU
template <typename DT>
static void banded_apply(u64 n, const DT *bands, const DT *rhs, const DT *x, DT *y) {
    std::vector<DenseVector<DT>> b;
    for (int k = 0; k < 9; ++k) {
        DenseVector<DT> v(n);
        for (u64 i = 0; i < n; ++i) v[i] = bands[(u64)k * n + i];
        b.push_back(v);
    }
    BandedMatrixQx<Q1Type, DT> A(n, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8]);
    DenseVector<DT> xv(n), yv(n, DT(0));
    for (u64 i = 0; i < n; ++i) xv[i] = x[i];
    if (rhs) {
        DenseVector<DT> rv(n);
        for (u64 i = 0; i < n; ++i) rv[i] = rhs[i];
        Defect<tags::CPU>::value(yv, rv, A, xv);
    } else {
        Product<tags::CPU>::value(yv, A, xv);
    }
    for (u64 i = 0; i < n; ++i) y[i] = yv[i];
}

extern "C" {

// ---- ELL handles ----------------------------------------------------------
void *ref_ell_from_sm_f(u64 rows, u64 cols, const u64 *rp, const u64 *col, const float *val, u64 threads) {
    set_threads(threads);
    return new SparseMatrixELL<float>(sm_from_csr(rows, cols, rp, col, val));
}
void *ref_ell_from_sm_d(u64 rows, u64 cols, const u64 *rp, const u64 *col, const double *val, u64 threads) {
    set_threads(threads);
    return new SparseMatrixELL<double>(sm_from_csr(rows, cols, rp, col, val));
}
void *ref_ell_from_csr_d(u64 rows, u64 cols, const u64 *rp, const u64 *col, const double *val, u64 threads) {
    set_threads(threads);
    Configuration::instance()->set_value("csr::blocksize", 1);
    SparseMatrix<double> sm(sm_from_csr(rows, cols, rp, col, val));
    SparseMatrixCSR<double> csr(sm);
    return new SparseMatrixELL<double>(csr);
}
void *ref_q1_ell_f(u64 side, int bc, u64 threads) { set_threads(threads); return q1_ell<float>(side, bc); }
void *ref_q1_ell_d(u64 side, int bc, u64 threads) { set_threads(threads); return q1_ell<double>(side, bc); }

void ref_ell_info_f(void *h, u64 *rows, u64 *cols, u64 *stride, u64 *ncpr, u64 *threads, u64 *size) {
    SparseMatrixELL<float> &m = *(SparseMatrixELL<float> *)h;
    *rows = m.rows(); *cols = m.columns(); *stride = m.stride(); *ncpr = m.num_cols_per_row();
    *threads = m.threads(); *size = m.Ax().size();
}
void ref_ell_info_d(void *h, u64 *rows, u64 *cols, u64 *stride, u64 *ncpr, u64 *threads, u64 *size) {
    SparseMatrixELL<double> &m = *(SparseMatrixELL<double> *)h;
    *rows = m.rows(); *cols = m.columns(); *stride = m.stride(); *ncpr = m.num_cols_per_row();
    *threads = m.threads(); *size = m.Ax().size();
}
void ref_ell_copy_f(void *h, u64 *Aj, float *Ax, u64 *Arl) {
    SparseMatrixELL<float> &m = *(SparseMatrixELL<float> *)h;
    for (u64 i = 0; i < m.Ax().size(); ++i) { Aj[i] = m.Aj()[i]; Ax[i] = m.Ax()[i]; }
    for (u64 i = 0; i < m.rows(); ++i) Arl[i] = m.Arl()[i];
}
void ref_ell_copy_d(void *h, u64 *Aj, double *Ax, u64 *Arl) {
    SparseMatrixELL<double> &m = *(SparseMatrixELL<double> *)h;
    for (u64 i = 0; i < m.Ax().size(); ++i) { Aj[i] = m.Aj()[i]; Ax[i] = m.Ax()[i]; }
    for (u64 i = 0; i < m.rows(); ++i) Arl[i] = m.Arl()[i];
}
void ref_ell_free_f(void *h) { delete (SparseMatrixELL<float> *)h; }
void ref_ell_free_d(void *h) { delete (SparseMatrixELL<double> *)h; }

void ref_banded_q1_f(u64 n, const float *bands, const float *rhs, const float *x, float *y) {
    banded_apply(n, bands, rhs, x, y);
}
void ref_banded_q1_d(u64 n, const double *bands, const double *rhs, const double *x, double *y) {
    banded_apply(n, bands, rhs, x, y);
}

// ---- operators on handles ---------------------------------------------------
void ref_product_f(void *h, const float *x, float *y) {
    SparseMatrixELL<float> &m = *(SparseMatrixELL<float> *)h;
    DenseVector<float> xv(m.columns()), yv(m.rows(), 0.f);
    for (u64 i = 0; i < m.columns(); ++i) xv[i] = x[i];
    Product<tags::CPU>::value(yv, m, xv);
    copy_out(yv, y);
}
void ref_product_d(void *h, const double *x, double *y) {
    SparseMatrixELL<double> &m = *(SparseMatrixELL<double> *)h;
    DenseVector<double> xv(m.columns()), yv(m.rows(), 0.);
    for (u64 i = 0; i < m.columns(); ++i) xv[i] = x[i];
    Product<tags::CPU>::value(yv, m, xv);
    copy_out(yv, y);
}
void ref_defect_d(void *h, const double *rhs, const double *x, double *r) {
    SparseMatrixELL<double> &m = *(SparseMatrixELL<double> *)h;
    DenseVector<double> xv(m.columns()), bv(m.rows()), rv(m.rows(), 0.);
    for (u64 i = 0; i < m.columns(); ++i) xv[i] = x[i];
    for (u64 i = 0; i < m.rows(); ++i) bv[i] = rhs[i];
    Defect<tags::CPU>::value(rv, bv, m, xv);
    copy_out(rv, r);
}
void ref_defect_f(void *h, const float *rhs, const float *x, float *r) {
    SparseMatrixELL<float> &m = *(SparseMatrixELL<float> *)h;
    DenseVector<float> xv(m.columns()), bv(m.rows()), rv(m.rows(), 0.f);
    for (u64 i = 0; i < m.columns(); ++i) xv[i] = x[i];
    for (u64 i = 0; i < m.rows(); ++i) bv[i] = rhs[i];
    Defect<tags::CPU>::value(rv, bv, m, xv);
    copy_out(rv, r);
}

// CGSolver<tags::CPU, NONE> (math/cg.hh:63-133)
u64 ref_cg_solver_d(void *h, const double *b, double *x, u64 max_iters, double eps) {
    SparseMatrixELL<double> &m = *(SparseMatrixELL<double> *)h;
    DenseVector<double> bv(m.rows()), xv(m.rows());
    for (u64 i = 0; i < m.rows(); ++i) { bv[i] = b[i]; xv[i] = x[i]; }
    u64 used = 0;
    unsigned long used_ul = 0;
    SparseMatrixELL<double> P(m);
    CGSolver<tags::CPU, methods::NONE>::value(m, P, bv, xv, (unsigned long)max_iters, used_ul, eps);
    used = used_ul;
    copy_out(xv, x);
    return used;
}

// ConjugateGradients<tags::CPU, JAC>::value(A, b, x, dd_inverted, max_iters,
// used_iters, eps) (math/conjugate_gradients.hh:1110-1154); its prints are
// swallowed.  Returns used_iters (0 when the reference leaves it unset).
extern "C++" template <typename DT>
u64 pcg_jacobi(void *h, const DT *b, DT *x, const DT *dinv, u64 max_iters, DT eps) {
    SparseMatrixELL<DT> &m = *(SparseMatrixELL<DT> *)h;
    DenseVector<DT> bv(m.rows()), xv(m.rows()), dv(m.rows());
    for (u64 i = 0; i < m.rows(); ++i) { bv[i] = b[i]; xv[i] = x[i]; dv[i] = dinv[i]; }
    unsigned long used_ul = 0;
    {
        CoutCapture cap;
        ConjugateGradients<tags::CPU, methods::JAC>::value(m, bv, xv, dv, (unsigned long)max_iters, used_ul, eps);
    }
    copy_out(xv, x);
    return used_ul;
}
u64 ref_pcg_jacobi_f(void *h, const float *b, float *x, const float *dinv, u64 max_iters, float eps) {
    return pcg_jacobi(h, b, x, dinv, max_iters, eps);
}
u64 ref_pcg_jacobi_d(void *h, const double *b, double *x, const double *dinv, u64 max_iters, double eps) {
    return pcg_jacobi(h, b, x, dinv, max_iters, eps);
}

// ConjugateGradients<tags::CPU, NONE>::value(A, b, x, konv_rad)
void ref_cg_konv_f(void *h, const float *b, float *x, double konv) {
    SparseMatrixELL<float> &m = *(SparseMatrixELL<float> *)h;
    DenseVector<float> bv(m.rows()), xv(m.rows(), 0.f);
    for (u64 i = 0; i < m.rows(); ++i) bv[i] = b[i];
    ConjugateGradients<tags::CPU, methods::NONE>::value(m, bv, xv, konv);
    copy_out(xv, x);
}

// ---- transfers ----------------------------------------------------------------
void ref_restriction_f(float *coarse, u64 nc, const float *fine, u64 nf, const u64 *mask) {
    DenseVector<float> c(nc, 0.f), f(nf);
    DenseVector<unsigned long> mk(8);
    for (u64 i = 0; i < nf; ++i) f[i] = fine[i];
    for (int i = 0; i < 8; ++i) mk[i] = mask[i];
    SparseMatrixELL<float> *dummy = nullptr;
    Restriction<tags::CPU, methods::NONE>::value(c, f, mk, dummy);
    copy_out(c, coarse);
}
void ref_prolongation_f(float *fine, u64 nf, const float *coarse, u64 nc, const u64 *mask) {
    DenseVector<float> c(nc), f(nf, 0.f);
    DenseVector<unsigned long> mk(8);
    for (u64 i = 0; i < nc; ++i) c[i] = coarse[i];
    for (int i = 0; i < 8; ++i) mk[i] = mask[i];
    SparseMatrixELL<float> *dummy = nullptr;
    Prolongation<tags::CPU, methods::NONE>::value(f, c, mk, dummy);
    copy_out(f, fine);
}
// ProlongationMatrix<tags::CPU>::value with natural ordering -> ELL handle
void *ref_prolmat_d(u64 fine_side, u64 threads) {
    set_threads(threads);
    const u64 cs = (fine_side + 1) / 2;
    SparseMatrix<double> P(fine_side * fine_side, cs * cs);
    DenseVector<unsigned long> fi(fine_side * fine_side), ci(cs * cs);
    for (u64 i = 0; i < fi.size(); ++i) fi[i] = i;
    for (u64 i = 0; i < ci.size(); ++i) ci[i] = i;
    ProlongationMatrix<tags::CPU>::value(P, fi, ci);
    return new SparseMatrixELL<double>(P);
}

// ---- Multigrid<tags::CPU, tags::CPU, NONE, JAC, V, MIXED> on the Dirichlet Q1
// problem of side 2^max_level+1, set up as math/poisson_mg_mixed_ell_TEST.cc
// does (mask all-Dirichlet here, initial_zero=false, adapt factor given).
// Returns outer iterations (parsed from the driver's own report).
u64 ref_mg_mixed(u64 max_level, u64 n_max_iter, u64 n_pre, double tol, double adapt,
                 const double *b, double *x, u64 *inner_total, double *solve_seconds) {
    set_threads(1);
    MGInfo<float, SparseMatrixELL<float> > info;
    info.is_smoother = false;
    for (int i = 0; i < 8; ++i) info.macro_border_mask[i] = 2;
    info.min_level = 1;
    info.max_level = max_level;
    info.n_max_iter = n_max_iter;
    info.initial_zero = false;
    info.tolerance = tol;
    info.convergence_check = false;
    info.n_pre_smooth = n_pre;
    info.n_post_smooth = n_pre;
    info.n_max_iter_coarse = 100;
    info.tolerance_coarse = 1e-2;
    info.adapt_correction_factor = adapt;
    // level 0 placeholders (never visited)
    {
        DenseVector<float> z(9, 0.f);
        SparseMatrixELL<float> *a0 = q1_ell<float>(3, 0);
        info.a.push_back(*a0);
        delete a0;
        info.c.push_back(z.copy()); info.d.push_back(z.copy());
        info.rhs.push_back(z.copy()); info.x.push_back(z.copy());
        info.diags_inverted.push_back(z.copy());
    }
    for (u64 l = 1; l <= max_level; ++l) {
        const u64 side = (1ul << l) + 1, n = side * side;
        DenseVector<float> z(n, 0.f);
        info.c.push_back(z.copy()); info.d.push_back(z.copy());
        info.x.push_back(z.copy()); info.rhs.push_back(z.copy());
        DenseVector<float> band(n);
        BandedMatrixQx<Q1Type, float> cm(n, band.copy(), band.copy(), band.copy(), band.copy(), band.copy(),
                                         band.copy(), band.copy(), band.copy(), band.copy());
        FillMatrix<tags::CPU, applications::POISSON, boundary_types::DIRICHLET::DIRICHLET_0>::value(cm);
        SparseMatrix<float> sm(cm);
        info.a.push_back(SparseMatrixELL<float>(sm));
        DenseVector<float> sdi(cm.band(DD).copy());
        ElementInverse<tags::CPU>::value(sdi);
        Scale<tags::CPU>::value(sdi, 0.7f);
        info.diags_inverted.push_back(sdi.copy());
    }
    const u64 side = (1ul << max_level) + 1, n = side * side;
    SparseMatrixELL<double> *A = q1_ell<double>(side, 0);
    DenseVector<double> rhs(n), res(n, 0.);
    for (u64 i = 0; i < n; ++i) { rhs[i] = b[i]; res[i] = x[i]; }
    u64 outer = 0;
    {
        CoutCapture cap;
        const auto t0 = std::chrono::steady_clock::now();
        Multigrid<tags::CPU, tags::CPU, methods::NONE, methods::JAC, methods::CYCLE::V, methods::MIXED>::value(
            *A, rhs, res, (unsigned long)11, std::numeric_limits<double>::epsilon(), info);
        const auto t1 = std::chrono::steady_clock::now();
        if (solve_seconds) *solve_seconds = std::chrono::duration<double>(t1 - t0).count();
        outer = cap.find_count("TN of outer iters: ");
        if (inner_total) *inner_total = cap.find_count("TN of inner iters: ");
    }
    copy_out(res, x);
    delete A;
    return outer;
}

}  // extern "C"
