// api_test.cpp -- tests of the C++ operator API on tags::GPU::HIP, in the
// style of HONEI's own *_TEST.cc files (checks against host computations,
// the reference's test patterns and its golden files).
//
//   api_test <golden-dir> [test-name-substring]
//
// Prints one "PASS <name>" / "FAIL <name>: <why>" line per test; exit status
// 0 when all selected tests pass.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <map>
#include <cstdlib>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "honei/backends/hip/memory_backend.hh"
#include "honei/la/algorithm.hh"
#include "honei/la/difference.hh"
#include "honei/la/dot_product.hh"
#include "honei/la/element_inverse.hh"
#include "honei/la/norm.hh"
#include "honei/la/product.hh"
#include "honei/la/scale.hh"
#include "honei/la/scaled_sum.hh"
#include "honei/la/sum.hh"
#include "honei/math/batch_solver.hh"
#include "honei/math/cg.hh"
#include "honei/math/defect.hh"
#include "honei/math/jacobi.hh"
#include "honei/math/mg.hh"
#include "honei/math/multigrid.hh"
#include "honei/math/prolongation.hh"
#include "honei/math/restriction.hh"

using namespace honei;

namespace {

std::string g_golden;

struct Failure {
    std::string what;
};

#define TEST_CHECK(c)                                                                  \
    do {                                                                               \
        if (!(c)) {                                                                    \
            std::ostringstream _o;                                                     \
            _o << __LINE__ << ": " #c;                                                 \
            throw Failure{_o.str()};                                                   \
        }                                                                              \
    } while (0)
#define TEST_CHECK_EQUAL_WITHIN_EPS(a, b, eps)                                         \
    do {                                                                               \
        const double _a = (double)(a), _b = (double)(b), _e = (double)(eps);           \
        if (!(std::fabs(_a - _b) <= _e)) {                                             \
            std::ostringstream _o;                                                     \
            _o << __LINE__ << ": |" #a " - " #b "| = " << std::fabs(_a - _b) << " > " << _e; \
            throw Failure{_o.str()};                                                   \
        }                                                                              \
    } while (0)

uint64_t g_state = 1;
double urand() {   // SplitMix64 -> U[-1, 1)
    g_state += 0x9E3779B97F4A7C15ull;
    uint64_t z = g_state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 2.0 * (double)(z >> 11) * (1.0 / 9007199254740992.0) - 1.0;
}

// la/sparse_matrix_ell_TEST.cc pattern: index % 5 == 0 -> index / 1.234 (+ 1)
template <typename DT>
SparseMatrix<DT> pattern(unsigned long rows, unsigned long cols, bool plus_one) {
    SparseMatrix<DT> s(rows, cols);
    for (unsigned long i = 0; i < rows * cols; ++i)
        if (i % 5 == 0) {
            const DT v = DT(double(i) / 1.234 + (plus_one ? 1.0 : 0.0));
            if (v != DT(0)) s(i / cols, i % cols, v);
        }
    return s;
}

// Q1 Dirichlet operator (math/fill_matrix.hh:36-78) as a SparseMatrix
template <typename DT>
SparseMatrix<DT> q1(unsigned long side) {
    SparseMatrix<DT> s(side * side, side * side);
    for (unsigned long i = 0; i < side; ++i)
        for (unsigned long j = 0; j < side; ++j) {
            const unsigned long r = i * side + j;
            if (i == 0 || j == 0 || i == side - 1 || j == side - 1) {
                s(r, r, DT(1.0));
                continue;
            }
            for (int di = -1; di <= 1; ++di)
                for (int dj = -1; dj <= 1; ++dj)
                    s(r, (i + di) * side + (j + dj), (di || dj) ? DT(-1.0 / 3.0) : DT(8.0 / 3.0));
        }
    return s;
}

template <typename DT>
std::vector<double> host_product(const SparseMatrix<DT> &s, const std::vector<double> &x) {
    std::vector<double> y(s.rows(), 0.0);
    for (unsigned long r = 0; r < s.rows(); ++r)
        for (auto &kv : s.row(r)) y[r] += (double)kv.second * x[kv.first];
    return y;
}

template <typename DT>
DenseVector<DT> vec(const std::vector<double> &v) {
    DenseVector<DT> r(v.size());
    for (size_t i = 0; i < v.size(); ++i) r.elements()[i] = (DT)v[i];
    return r;
}

std::vector<double> rand_vec(unsigned long n) {
    std::vector<double> v(n);
    for (auto &e : v) e = urand();
    return v;
}

// ---------------------------------------------------------------------------
template <typename DT>
void ell_quick() {
    for (unsigned long threads = 1; threads <= 16; threads *= 2) {
        Configuration::instance()->set_value("ell::threads", (int)threads);
        SparseMatrix<DT> sms = pattern<DT>(111, 114, true);
        sms(0, 0, DT(4711));
        SparseMatrixELL<DT> sm0(sms);
        TEST_CHECK(sm0 == sm0);
        TEST_CHECK(sm0 == sm0.copy());
        TEST_CHECK(sms.used_elements() == sm0.used_elements());
        TEST_CHECK(sm0.stride() == 32 * ((111 * threads + 31) / 32));
        for (unsigned long r = 0; r < 111; ++r)
            for (auto &kv : sms.row(r)) TEST_CHECK(sm0(r, kv.first) == kv.second);
    }
    Configuration::instance()->set_value("ell::threads", 1);
}

template <typename DT>
void product_ell() {
    // la/product_TEST.cc:786-803
    for (unsigned long size = 11; size < (1 << 9); size <<= 1) {
        SparseMatrix<DT> sms = pattern<DT>(size, size + 3, false);
        std::vector<double> x(size + 3, 4.0);
        x[0] = 1;
        x[1] = 2;
        const std::vector<double> ref = host_product(sms, x);
        for (unsigned long threads = 1; threads <= 16; threads *= 2) {
            Configuration::instance()->set_value("ell::threads", (int)threads);
            SparseMatrixELL<DT> sm0(sms);
            DenseVector<DT> dv1 = vec<DT>(x);
            DenseVector<DT> prod(size, DT(4711));
            Product<tags::GPU::HIP>::value(prod, sm0, dv1);
            const double eps = sizeof(DT) == 4 ? 1e-6 : 1e-13;
            for (unsigned long i = 0; i < size; ++i) {
                double scale = 0;
                for (auto &kv : sms.row(i)) scale += std::fabs((double)kv.second * x[kv.first]);
                TEST_CHECK_EQUAL_WITHIN_EPS(prod[i], ref[i], eps * scale + 1e-30);
            }
            // x staged in LDS: the same bits
            SparseMatrixELL<DT> sm1(sms);
            sm1.set_stage_x(true);
            DenseVector<DT> prod_s(size, DT(4711));
            Product<tags::GPU::HIP>::value(prod_s, sm1, dv1);
            TEST_CHECK(sm1.device().window > 0);
            for (unsigned long i = 0; i < size; ++i) TEST_CHECK(prod_s[i] == prod[i]);
        }
    }
    Configuration::instance()->set_value("ell::threads", 1);
}

// MemoryBackend<tags::GPU::HIP> against the contract of
// util/memory_backend_base.hh:32-58 as the arbiter uses it
// (util/memory_arbiter-impl.hh:380-500): upload of a known block returns its
// copy without transferring, download writes back to the block's host
// address, free forgets it; copy, fill and both conversions act on the
// device copies.
void memory_backend_api() {
    MemoryBackend<tags::GPU::HIP> *mb = MemoryBackend<tags::GPU::HIP>::instance();
    const unsigned long n = 1000;
    std::vector<double> a(n), b(n, 0.0), c(n, -1.0);
    std::vector<float> f(n);
    for (unsigned long i = 0; i < n; ++i) a[i] = urand();
    const unsigned long blocks0 = mb->blocks();
    TEST_CHECK(!mb->knows(a.data(), a.data()));
    void *da = mb->upload(a.data(), a.data(), 8 * n);
    TEST_CHECK(mb->knows(a.data(), a.data()) && mb->blocks() == blocks0 + 1);
    const double keep = a[7];
    a[7] = 12345.0;   // host change without a new lock: the known block is not re-sent
    TEST_CHECK(mb->upload(a.data(), a.data(), 8 * n) == da);
    mb->download(a.data(), a.data(), 8 * n);
    TEST_CHECK(a[7] == keep);
    void *db = mb->alloc(b.data(), b.data(), 8 * n);
    TEST_CHECK(mb->alloc(b.data(), b.data(), 8 * n) == db);
    mb->copy(a.data(), a.data(), b.data(), b.data(), 8 * n);
    mb->download(b.data(), b.data(), 8 * n);
    for (unsigned long i = 0; i < n; ++i) TEST_CHECK(b[i] == a[i]);
    mb->alloc(f.data(), f.data(), 4 * n);
    mb->convert_double_float(a.data(), a.data(), f.data(), f.data(), 8 * n);
    mb->download(f.data(), f.data(), 4 * n);
    for (unsigned long i = 0; i < n; ++i) TEST_CHECK(f[i] == (float)a[i]);
    mb->alloc(c.data(), c.data(), 8 * n);
    mb->convert_float_double(f.data(), f.data(), c.data(), c.data(), 4 * n);
    mb->download(c.data(), c.data(), 8 * n);
    for (unsigned long i = 0; i < n; ++i) TEST_CHECK(c[i] == (double)f[i]);
    mb->fill(c.data(), c.data(), 8 * n, 2.5);
    mb->fill(f.data(), f.data(), 4 * n, -0.75f);
    mb->download(c.data(), c.data(), 8 * n);
    mb->download(f.data(), f.data(), 4 * n);
    for (unsigned long i = 0; i < n; ++i) TEST_CHECK(c[i] == 2.5 && f[i] == -0.75f);
    for (void *id : {(void *)a.data(), (void *)b.data(), (void *)c.data(), (void *)f.data()}) mb->free(id);
    TEST_CHECK(!mb->knows(a.data(), a.data()) && mb->blocks() == blocks0);
    bool threw = false;
    try {
        mb->download(a.data(), a.data(), 8 * n);
    } catch (InternalError &) {
        threw = true;
    }
    TEST_CHECK(threw);
}

// The reference's glue as written (la/product-cuda.cc:178-194,
// math/defect-cuda.cc:524-660): the matrix's Aj, Ax, Arl and the vectors are
// locked through the memory backend -- Aj and Arl as the 64-bit unsigned
// long blocks SparseMatrixELL holds -- and handed to the launcher.  The
// *_idx64 launchers give the bits of the int32 path (Product / Defect
// <tags::GPU::HIP> on the same matrix), and the device-side narrowing of
// those blocks gives the int32 arrays.
template <typename DT>
void reference_glue() {
    MemoryBackend<tags::GPU::HIP> *mb = MemoryBackend<tags::GPU::HIP>::instance();
    auto product64 = sizeof(DT) == 4 ? honei_hip_product_smell_dv_float_idx64 : honei_hip_product_smell_dv_double_idx64;
    auto defect64 = sizeof(DT) == 4 ? honei_hip_defect_smell_dv_float_idx64 : honei_hip_defect_smell_dv_double_idx64;
    for (unsigned long threads = 1; threads <= 16; threads *= 2) {
        Configuration::instance()->set_value("ell::threads", (int)threads);
        for (int which = 0; which < 2; ++which) {
            SparseMatrix<DT> sms = which == 0 ? pattern<DT>(300, 303, true) : q1<DT>(33);
            SparseMatrixELL<DT> a(sms);
            const unsigned long rows = a.rows(), cols = a.columns(), slots = a.Ax().size();
            DenseVector<DT> x = vec<DT>(rand_vec(cols)), rhs = vec<DT>(rand_vec(rows));
            DenseVector<DT> y32(rows, DT(0)), r32(rows, DT(0));
            Product<tags::GPU::HIP>::value(y32, a, x);
            // Defect<> keeps the reference's size checks (rhs against the
            // columns, math/defect-cuda.cc:530-537): square matrices only
            const bool square = rows == cols;
            if (square) Defect<tags::GPU::HIP>::value(r32, rhs, a, x);
            unsigned long *aj = a.Aj().elements(), *arl = a.Arl().elements();
            DT *ax = a.Ax().elements(), *xe = x.elements(), *re = rhs.elements();
            std::vector<DT> y(rows, DT(4711)), r(rows, DT(4711));
            void *Aj_gpu = mb->upload(aj, aj, 8 * slots);
            void *Ax_gpu = mb->upload(ax, ax, sizeof(DT) * slots);
            void *Arl_gpu = mb->upload(arl, arl, 8 * rows);
            void *x_gpu = mb->upload(xe, xe, sizeof(DT) * cols);
            void *rhs_gpu = mb->upload(re, re, sizeof(DT) * rows);
            void *y_gpu = mb->alloc(y.data(), y.data(), sizeof(DT) * rows);
            void *r_gpu = mb->alloc(r.data(), r.data(), sizeof(DT) * rows);
            hip_check(product64(x_gpu, y_gpu, Aj_gpu, Ax_gpu, Arl_gpu, 0, rows, a.num_cols_per_row(), a.stride(), 0,
                                a.threads(), nullptr),
                      "product idx64");
            if (square)
                hip_check(defect64(rhs_gpu, r_gpu, Aj_gpu, Ax_gpu, Arl_gpu, x_gpu, 0, rows, a.num_cols_per_row(),
                                   a.stride(), 0, a.threads(), nullptr),
                          "defect idx64");
            mb->download(y.data(), y.data(), sizeof(DT) * rows);
            mb->download(r.data(), r.data(), sizeof(DT) * rows);
            for (unsigned long i = 0; i < rows; ++i) TEST_CHECK(y[i] == y32[i] && (!square || r[i] == r32[i]));
            // a sub-range writes y[i - row_start], as the int32 entry does
            const unsigned long r0 = rows / 3, r1 = rows - 5;
            hip_check(product64(x_gpu, y_gpu, Aj_gpu, Ax_gpu, Arl_gpu, r0, r1, a.num_cols_per_row(), a.stride(), 0,
                                a.threads(), nullptr),
                      "product idx64 range");
            mb->download(y.data(), y.data(), sizeof(DT) * rows);
            for (unsigned long i = r0; i < r1; ++i) TEST_CHECK(y[i - r0] == y32[i]);
            // device-side narrowing of the 64-bit blocks
            std::vector<uint32_t> aj32(slots), arl32(rows);
            void *d32 = nullptr;
            hip_check(honei_hip_malloc(&d32, 4 * slots + 4 * rows + 8), "malloc");
            hip_check(honei_hip_narrow_indices(d32, Aj_gpu, slots, cols, nullptr), "narrow Aj");
            hip_check(honei_hip_download(aj32.data(), d32, 4 * slots, nullptr), "download");
            hip_check(honei_hip_narrow_indices(d32, Arl_gpu, rows, a.num_cols_per_row() + 1, nullptr), "narrow Arl");
            hip_check(honei_hip_download(arl32.data(), d32, 4 * rows, nullptr), "download");
            for (unsigned long i = 0; i < slots; ++i) TEST_CHECK(aj32[i] == aj[i]);
            for (unsigned long i = 0; i < rows; ++i) TEST_CHECK(arl32[i] == arl[i]);
            // an index at or above the bound is reported, not narrowed silently
            TEST_CHECK(honei_hip_narrow_indices(d32, Aj_gpu, slots, 1, nullptr) == HONEI_HIP_ERROR_INVALID_ARGUMENT);
            hip_check(honei_hip_free(d32), "free");
            for (void *id : {(void *)aj, (void *)ax, (void *)arl, (void *)xe, (void *)re, (void *)y.data(),
                             (void *)r.data()})
                mb->free(id);
        }
    }
    Configuration::instance()->set_value("ell::threads", 1);
}

template <typename DT>
void defect_and_blas1() {
    const unsigned long side = 65, n = side * side;
    SparseMatrix<DT> s = q1<DT>(side);
    SparseMatrixELL<DT> A(s);
    const std::vector<double> xv = rand_vec(n), bv = rand_vec(n), cv = rand_vec(n);
    DenseVector<DT> x = vec<DT>(xv), b = vec<DT>(bv), c = vec<DT>(cv);
    const double eps = sizeof(DT) == 4 ? 1e-5 : 1e-13;
    DenseVector<DT> r(n);
    Defect<tags::GPU::HIP>::value(r, b, A, x);
    DenseVector<DT> r2 = Defect<tags::GPU::HIP>::value(b, A, x);
    std::vector<double> xr(n);
    for (unsigned long i = 0; i < n; ++i) xr[i] = (double)x[i];
    const std::vector<double> ax = host_product(s, xr);
    for (unsigned long i = 0; i < n; ++i) {
        TEST_CHECK_EQUAL_WITHIN_EPS(r[i], (double)b[i] - ax[i], eps * 4);
        TEST_CHECK(r[i] == r2[i]);
    }
    double nn = 0, dd = 0;
    for (unsigned long i = 0; i < n; ++i) {
        nn += (double)x[i] * (double)x[i];
        dd += (double)x[i] * (double)b[i];
    }
    TEST_CHECK_EQUAL_WITHIN_EPS((Norm<vnt_l_two, false, tags::GPU::HIP>::value(x)), nn, eps * nn);
    TEST_CHECK_EQUAL_WITHIN_EPS((Norm<vnt_l_two, true, tags::GPU::HIP>::value(x)), std::sqrt(nn), eps * std::sqrt(nn));
    TEST_CHECK_EQUAL_WITHIN_EPS(DotProduct<tags::GPU::HIP>::value(x, b), dd, eps * n);
    // ScaledSum three forms, Sum, Difference, Scale, ElementProduct, ElementInverse
    DenseVector<DT> t = x.copy();
    ScaledSum<tags::GPU::HIP>::value(t, b, DT(0.5));
    for (unsigned long i = 0; i < n; i += 97) TEST_CHECK(t[i] == x[i] + b[i] * DT(0.5));
    DenseVector<DT> u(n);
    ScaledSum<tags::GPU::HIP>::value(u, x, b, DT(-2));
    for (unsigned long i = 0; i < n; i += 97) TEST_CHECK(u[i] == x[i] + b[i] * DT(-2));
    DenseVector<DT> w = x.copy();
    ScaledSum<tags::GPU::HIP>::value(w, b, c);
    for (unsigned long i = 0; i < n; i += 97) TEST_CHECK(w[i] == x[i] + b[i] * c[i]);
    DenseVector<DT> z = x.copy();
    Sum<tags::GPU::HIP>::value(z, b);
    for (unsigned long i = 0; i < n; i += 97) TEST_CHECK(z[i] == x[i] + b[i]);
    Difference<tags::GPU::HIP>::value(z, b);
    DenseVector<DT> z3(n);
    Difference<tags::GPU::HIP>::value(z3, x, b);
    Scale<tags::GPU::HIP>::value(z3, DT(3));
    DenseVector<DT> e(n);
    ElementProduct<tags::GPU::HIP>::value(e, x, b);
    DenseVector<DT> inv(n, DT(4));
    ElementInverse<tags::GPU::HIP>::value(inv);
    for (unsigned long i = 0; i < n; i += 97) {
        TEST_CHECK(z[i] == (x[i] + b[i]) - b[i]);
        TEST_CHECK(z3[i] == (x[i] - b[i]) * DT(3));
        TEST_CHECK(e[i] == x[i] * b[i]);
        TEST_CHECK(inv[i] == DT(0.25));
    }
    // copy / fill / convert
    DenseVector<DT> cp(n);
    copy<tags::GPU::HIP>(x, cp);
    fill<tags::GPU::HIP>(z, DT(7));
    for (unsigned long i = 0; i < n; i += 97) TEST_CHECK(cp[i] == x[i] && z[i] == DT(7));
    DenseVector<float> f(n);
    DenseVector<double> dbl = vec<double>(xv);
    convert<tags::GPU::HIP>(f, dbl);
    DenseVector<double> back(n);
    convert<tags::GPU::HIP>(back, f);
    for (unsigned long i = 0; i < n; i += 97) TEST_CHECK(back[i] == (double)(float)xv[i]);
    // size mismatch throws like the reference's glue
    bool thrown = false;
    try {
        DenseVector<DT> small(n - 1);
        Product<tags::GPU::HIP>::value(small, A, x);
    } catch (VectorSizeDoesNotMatch &) {
        thrown = true;
    }
    TEST_CHECK(thrown);
}

template <typename DT>
void jacobi() {
    const unsigned long side = 33, n = side * side;
    SparseMatrix<DT> s = q1<DT>(side);
    SparseMatrixELL<DT> A(s);
    DenseVector<DT> dinv(n);
    for (unsigned long i = 0; i < n; ++i) dinv[i] = (DT(1) / A(i, i)) * DT(0.7);
    const std::vector<double> bv = rand_vec(n);
    DenseVector<DT> b = vec<DT>(bv);
    DenseVector<DT> x(n);
    Jacobi<tags::GPU::HIP>::value(A, b, x, DT(0.7), dinv);
    std::vector<double> h(n);
    for (unsigned long i = 0; i < n; ++i) {
        TEST_CHECK(x[i] == dinv[i] * b[i]);
        h[i] = (double)x[i];
    }
    Jacobi<tags::GPU::HIP>::value(x, A, b, x, 3, DT(0.7), dinv);
    for (int k = 0; k < 3; ++k) {
        const std::vector<double> ah = host_product(s, h);
        std::vector<double> hn(n);
        for (unsigned long i = 0; i < n; ++i) hn[i] = h[i] + (double)dinv[i] * ((double)b[i] - ah[i]);
        h = hn;
    }
    const double eps = sizeof(DT) == 4 ? 1e-5 : 1e-13;
    for (unsigned long i = 0; i < n; ++i) TEST_CHECK_EQUAL_WITHIN_EPS(x[i], h[i], eps);
}

template <typename DT>
void transfers() {
    // geometric == PROLMAT + Dirichlet (math/prolongation_TEST.cc:93-138)
    for (unsigned long level = 1; level <= 6; ++level) {
        const unsigned long sc = (1ul << level) + 1, sf = 2 * sc - 1;
        SparseMatrix<DT> P(sf * sf, sc * sc);
        for (unsigned long i = 0; i < sf; ++i)
            for (unsigned long j = 0; j < sf; ++j)
                for (unsigned long ci = 0; ci < sc; ++ci)
                    for (unsigned long cj = 0; cj < sc; ++cj) {
                        const double wi = std::max(0.0, 1.0 - std::fabs(double(i) - 2.0 * ci) / 2.0);
                        const double wj = std::max(0.0, 1.0 - std::fabs(double(j) - 2.0 * cj) / 2.0);
                        if (wi * wj > 0) P(i * sf + j, ci * sc + cj, DT(wi * wj));
                    }
        SparseMatrixELL<DT> Pe(P);
        DenseVector<DT> coarse = vec<DT>(rand_vec(sc * sc));
        DenseVector<DT> f1(sf * sf), f2(sf * sf);
        DenseVector<unsigned long> mask(8, 2ul);
        Prolongation<tags::GPU::HIP, methods::NONE>::value(f1, coarse, mask, Pe);
        Prolongation<tags::GPU::HIP, methods::PROLMAT>::value(f2, coarse, mask, Pe);
        for (unsigned long i = 0; i < sf; ++i)
            for (unsigned long j = 0; j < sf; ++j) {
                const bool bd = i == 0 || j == 0 || i == sf - 1 || j == sf - 1;
                TEST_CHECK_EQUAL_WITHIN_EPS(f1[i * sf + j], bd ? 0.0 : (double)f2[i * sf + j], 1e-6);
            }
        // restriction of a constant field: 4 inside, 0 on Dirichlet nodes
        DenseVector<DT> ones(sf * sf, DT(1)), rc(sc * sc);
        Restriction<tags::GPU::HIP, methods::NONE>::value(rc, ones, mask, Pe);
        for (unsigned long i = 0; i < sc; ++i)
            for (unsigned long j = 0; j < sc; ++j) {
                const bool bd = i == 0 || j == 0 || i == sc - 1 || j == sc - 1;
                TEST_CHECK(rc[i * sc + j] == (bd ? DT(0) : DT(4)));
            }
    }
}

void cg_solver() {
    const unsigned long side = 129, n = side * side;
    SparseMatrix<double> s = q1<double>(side);
    SparseMatrixELL<double> A(s);
    const std::vector<double> xt = rand_vec(n);
    const std::vector<double> bh = host_product(s, xt);
    DenseVector<double> b = vec<double>(bh), x(n, 0.0);
    unsigned long used = 0;
    CGSolver<tags::GPU::HIP, methods::NONE>::value(A, A, b, x, 10000ul, used, 1e-8);
    TEST_CHECK(used > 0 && used < 10000);
    DenseVector<double> r = Defect<tags::GPU::HIP>::value(b, A, x);
    const double rn = Norm<vnt_l_two, true, tags::GPU::HIP>::value(r);
    const double bn = Norm<vnt_l_two, true, tags::GPU::HIP>::value(b);
    TEST_CHECK(rn < 1.01e-8 * bn);
    double err = 0;
    for (unsigned long i = 0; i < n; ++i) err = std::max(err, std::fabs(x[i] - xt[i]));
    TEST_CHECK(err < 1e-5);
    // ConjugateGradients<NONE>::value(A, b, x, max_iters): the same iteration
    // with the fixed 1e-8 tolerance, so the same result
    DenseVector<double> x2(n, 0.0);
    ConjugateGradients<tags::GPU::HIP, methods::NONE>::value(A, b, x2, 10000ul);
    for (unsigned long i = 0; i < n; ++i) TEST_CHECK(x2[i] == x[i]);
}

// math/dune_regression_TEST.cc:66-178: PCG-Jacobi on the area51 system from
// its start vector, checked against the FEATFLOW2 solution with that test's
// tolerance (3e-5).
void pcg_jacobi_area51() {
    const std::string file = g_golden + "/l2/area51_";
    SparseMatrixELL<double> A(MatrixIO<io_formats::ELL>::read_matrix(file + "full_0.ell", double(0)));
    DenseVector<double> rhs(VectorIO<io_formats::EXP>::read_vector(file + "rhs_0", double(0)));
    DenseVector<double> x(VectorIO<io_formats::EXP>::read_vector(file + "init_0", double(0)));
    DenseVector<double> sol(VectorIO<io_formats::EXP>::read_vector(file + "sol_0", double(0)));
    const unsigned long n = A.rows();
    DenseVector<double> dinv(n, 0.0);
    for (unsigned long r = 0; r < n; ++r)
        for (unsigned long k = 0; k < A.Arl()[r]; ++k) {
            const unsigned long slot = r + k * A.stride();
            if (A.Aj()[slot] == r) dinv[r] = 1.0 / A.Ax()[slot];
        }
    unsigned long used = 0;
    ConjugateGradients<tags::GPU::HIP, methods::JAC>::value(A, rhs, x, dinv, 10000ul, used, 1e-8);
    TEST_CHECK(used > 0 && used < 10000);
    for (unsigned long i = 0; i < n; ++i) TEST_CHECK_EQUAL_WITHIN_EPS(x[i], sol[i], 3e-5);
}

// math/poisson_mg_mixed_ell_TEST.cc:42-230 set-up, Dirichlet on all edges
void multigrid_mixed() {
    const unsigned long max_level = 6, side = (1ul << max_level) + 1, n = side * side;
    MGInfo<float, SparseMatrixELL<float>> info;
    info.min_level = 1;
    info.max_level = max_level;
    info.n_max_iter = 4;
    info.tolerance = 1e-8f;
    info.n_pre_smooth = info.n_post_smooth = 2;
    info.adapt_correction_factor = 1.f;
    for (unsigned long l = 0; l <= max_level; ++l) {
        const unsigned long sl = (1ul << l) + 1, nl = sl * sl;
        SparseMatrixELL<float> A(q1<float>(l == 0 ? 3 : sl));
        DenseVector<float> dinv(A.rows());
        for (unsigned long i = 0; i < A.rows(); ++i) dinv[i] = (1.f / A(i, i)) * 0.7f;
        info.a.push_back(A);
        info.diags_inverted.push_back(dinv);
        info.c.push_back(DenseVector<float>(nl, 0.f));
        info.d.push_back(DenseVector<float>(nl, 0.f));
        info.rhs.push_back(DenseVector<float>(nl, 0.f));
        info.x.push_back(DenseVector<float>(nl, 0.f));
    }
    SparseMatrix<double> s = q1<double>(side);
    SparseMatrixELL<double> system(s);
    const std::vector<double> xt = rand_vec(n);
    DenseVector<double> rhs = vec<double>(host_product(s, xt)), result(n, 0.0);
    Multigrid<tags::GPU::HIP, tags::GPU::HIP, methods::NONE, methods::JAC, methods::CYCLE::V, methods::MIXED>::value(
        system, rhs, result, 11ul, std::numeric_limits<double>::epsilon(), info);
    // the reference settles at 3 outer steps of 4 V-cycles for this set-up at every level
    TEST_CHECK(info.outer_iterations >= 2 && info.outer_iterations <= 4);
    TEST_CHECK(info.inner_iterations == 4 * info.outer_iterations);
    double err = 0;
    for (unsigned long i = 0; i < n; ++i) err = std::max(err, std::fabs(result[i] - xt[i]));
    TEST_CHECK(err < 1e-6);
}

// math/mg_TEST.cc:389-464: W-cycle, CG coarse, Richardson/Jacobi 0.7, 4+4
// sweeps, PROLMAT, testdata/poisson levels 1..6 against FEAT's sol_6.
void mg_solver_quad() {
    const unsigned long max_level = 6, min_level = 1;
    const std::string file = g_golden + "/poisson/";
    typedef MGUtil<tags::GPU::HIP, SparseMatrixELL<double>, DenseVector<double>, SparseMatrixELL<double>,
                   DenseVector<double>, MatrixIO<io_formats::ELL>, VectorIO<io_formats::EXP>, double>
        Util;
    auto data = Util::load_data(file, max_level, 0.7, "jac");
    Util::configure(data, 100, 100, 4, 4, min_level, 1e-8);
    OperatorList ol(MGCycleCreation<tags::GPU::HIP, methods::CYCLE::W::STATIC, CGSolver<tags::GPU::HIP, methods::NONE>,
                                    RISmoother<tags::GPU::HIP>, Restriction<tags::GPU::HIP, methods::PROLMAT>,
                                    Prolongation<tags::GPU::HIP, methods::PROLMAT>, double>::value(data));
    MGSolver<tags::GPU::HIP, Norm<vnt_l_two, true, tags::GPU::HIP>>::value(data, ol);
    DenseVector<double> ref(VectorIO<io_formats::EXP>::read_vector(file + "sol_6", double(0)));
    const double base_eps = 1e-1, add_eps = 1e-3;
    const double m = (add_eps - base_eps) / 4.0, bb = base_eps - 4.0 * m;
    const double eps = (m * sizeof(double) + bb) * 8.0;
    DenseVector<double> &x = data.x.at(MGDataIndex::internal_index_A(max_level));
    TEST_CHECK(data.used_iters > 0 && data.used_iters < 100);
    for (unsigned long i = 0; i < ref.size(); ++i) TEST_CHECK_EQUAL_WITHIN_EPS(x[i], ref[i], eps);
    double err = 0;
    for (unsigned long i = 0; i < ref.size(); ++i) err = std::max(err, std::fabs(x[i] - ref[i]));
    TEST_CHECK(err < 1e-5);
}

// ---- batched multigrid and the request loop (no reference counterpart) -------

// Per-cell coefficient exp(sin(i) + cos(3 j)) (smooth, contrast ~e^2) and the
// FP64 variable-coefficient Q1 operator assembled on the host from element
// stiffness matrices (independent of the device assembly).
static DenseVector<double> coefficient(unsigned long side, double shift) {
    const unsigned long nc = side - 1;
    DenseVector<double> k(nc * nc);
    for (unsigned long i = 0; i < nc; ++i)
        for (unsigned long j = 0; j < nc; ++j) k[i * nc + j] = std::exp(std::sin(0.7 * i + shift) + std::cos(3.1 * j));
    return k;
}

static SparseMatrixELL<double> host_operator(unsigned long side, const DenseVector<double> &k) {
    const double K[4][4] = {{4, -1, -2, -1}, {-1, 4, -1, -2}, {-2, -1, 4, -1}, {-1, -2, -1, 4}};
    const unsigned long n = side * side, nc = side - 1;
    std::vector<std::map<unsigned long, double>> rows(n);
    for (unsigned long ci = 0; ci < nc; ++ci)
        for (unsigned long cj = 0; cj < nc; ++cj) {
            const unsigned long nd[4] = {ci * side + cj, ci * side + cj + 1, (ci + 1) * side + cj + 1,
                                         (ci + 1) * side + cj};
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) rows[nd[a]][nd[b]] += k[ci * nc + cj] * K[a][b] / 6.0;
        }
    SparseMatrix<double> sm(n, n);
    for (unsigned long r = 0; r < n; ++r) {
        const unsigned long i = r / side, j = r % side;
        if (i == 0 || j == 0 || i == side - 1 || j == side - 1) {
            sm(r, r, 1.0);
            continue;
        }
        for (auto &kv : rows[r]) sm(r, kv.first, kv.second);
    }
    return SparseMatrixELL<double>(sm);
}

static void batch_solver() {
    const std::vector<unsigned long> levels = {3, 5, 4};
    BatchConfig cfg;
    cfg.autotune = 0;
    BatchSolver<tags::GPU::HIP> solver(levels, {1, 2, 1}, cfg);
    TEST_CHECK(solver.slots() == 4 && solver.device_bytes() > 0);
    struct Sys {
        unsigned long level, slot;
        DenseVector<double> k, b;
    };
    std::vector<Sys> sys;
    const unsigned long js[3] = {1, 2, 1};
    for (unsigned c = 0; c < 3; ++c)
        for (unsigned long j = 0; j < js[c]; ++j) {
            const unsigned long side = (1ul << levels[c]) + 1;
            DenseVector<double> b(side * side);
            for (unsigned long i = 0; i < side * side; ++i) b[i] = std::sin(0.37 * i + c + j);
            sys.push_back({levels[c], solver.slot(c, j), coefficient(side, 0.3 * c + j), b});
        }
    for (auto &s : sys) solver.load(s.slot, s.k, s.b);
    const unsigned long launches = solver.run();
    TEST_CHECK(launches > 0);
    std::vector<SlotInfo> info = solver.poll();
    for (auto &s : sys) {
        const SlotInfo &in = info[s.slot];
        TEST_CHECK(in.state == 0 && in.status == 0 && in.level == (int32_t)s.level && in.rel < 1e-8);
        const unsigned long side = (1ul << s.level) + 1, n = side * side;
        DenseVector<double> x(n);
        solver.result(s.slot, x);
        // FP64 residual with the host-assembled operator, against the start
        // vector's (first row, last row, first column = b)
        SparseMatrixELL<double> A = host_operator(side, s.k);
        DenseVector<double> x0(n, 0.0);
        for (unsigned long i = 0; i < n; ++i)
            if (i < side || i >= n - side || i % side == 0) x0[i] = s.b[i];
        auto rnorm = [&](const DenseVector<double> &v) {
            double acc = 0.0;
            for (unsigned long r = 0; r < n; ++r) {
                double y = 0.0;
                for (unsigned long c = 0; c < n; ++c) {
                    const double a = A(r, c);
                    if (a != 0.0) y += a * v[c];
                }
                acc += (s.b[r] - y) * (s.b[r] - y);
            }
            return std::sqrt(acc);
        };
        TEST_CHECK(rnorm(x) / rnorm(x0) < 1.001e-8);
    }
    // the same system alone in a fresh pool: bit-identical solution
    BatchSolver<tags::GPU::HIP> alone({5}, {1}, cfg);
    alone.load(alone.slot(0, 0), sys[1].k, sys[1].b);
    alone.run();
    DenseVector<double> xa(33 * 33), xm(33 * 33);
    alone.result(alone.slot(0, 0), xa);
    solver.result(sys[1].slot, xm);
    for (unsigned long i = 0; i < 33 * 33; ++i) TEST_CHECK(xa[i] == xm[i]);
}

static void serve_loop() {
    honei_hip_serve_config cfg{};
    hip_check(honei_hip_batch_config_default(&cfg.batch), "config");
    cfg.nclass = 2;
    cfg.levels[0] = 4;
    cfg.levels[1] = 5;
    cfg.counts[0] = 2;
    cfg.counts[1] = 1;
    cfg.sigma = 1.0;
    std::vector<honei_hip_request> reqs;
    for (int i = 0; i < 9; ++i) {
        honei_hip_request r{};
        r.arrival_s = 0.001 * i;
        r.level = (i % 3 == 2) ? 5 : 4;
        r.coeff_seed = 100 + 2 * i;
        r.rhs_seed = 101 + 2 * i;
        reqs.push_back(r);
    }
    std::vector<honei_hip_request_result> res;
    const honei_hip_serve_stats st = ServeLoop::run(cfg, reqs, res);
    TEST_CHECK(st.completed == 9 && st.rejected == 0 && st.not_converged == 0);
    for (size_t i = 0; i < res.size(); ++i) {
        TEST_CHECK(res[i].status == 0 && res[i].rel < 1e-8);
        TEST_CHECK(res[i].admit_s >= reqs[i].arrival_s && res[i].done_s >= res[i].admit_s);
    }
    TEST_CHECK(st.p50_ms <= st.p99_ms && st.p99_ms <= st.max_ms);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <golden-dir> [filter]\n", argv[0]);
        return 2;
    }
    g_golden = argv[1];
    const std::string filter = argc > 2 ? argv[2] : "";
    if (honei_hip_init(0) != HONEI_HIP_SUCCESS) {
        std::printf("FAIL init: %s\n", honei_hip_last_error());
        return 1;
    }
    std::vector<std::pair<std::string, std::function<void()>>> tests = {
        {"sparse_matrix_ell_quick_test<float>", ell_quick<float>},
        {"sparse_matrix_ell_quick_test<double>", ell_quick<double>},
        {"sparse_matrix_ell_dense_vector_product_test<float>", product_ell<float>},
        {"sparse_matrix_ell_dense_vector_product_test<double>", product_ell<double>},
        {"memory_backend_api", memory_backend_api},
        {"reference_glue_idx64<float>", reference_glue<float>},
        {"reference_glue_idx64<double>", reference_glue<double>},
        {"defect_and_blas1<float>", defect_and_blas1<float>},
        {"defect_and_blas1<double>", defect_and_blas1<double>},
        {"jacobi<float>", jacobi<float>},
        {"jacobi<double>", jacobi<double>},
        {"transfers<float>", transfers<float>},
        {"transfers<double>", transfers<double>},
        {"cg_solver<double>", cg_solver},
        {"pcg_jacobi_area51<double>", pcg_jacobi_area51},
        {"multigrid_mixed", multigrid_mixed},
        {"mg_solver_quad_w_cycle", mg_solver_quad},
        {"batch_solver", batch_solver},
        {"serve_loop", serve_loop},
    };
    int failed = 0, run = 0;
    for (auto &t : tests) {
        if (!filter.empty() && t.first.find(filter) == std::string::npos) continue;
        ++run;
        try {
            t.second();
            std::printf("PASS %s\n", t.first.c_str());
        } catch (Failure &f) {
            ++failed;
            std::printf("FAIL %s: %s\n", t.first.c_str(), f.what.c_str());
        } catch (std::exception &e) {
            ++failed;
            std::printf("FAIL %s: exception %s\n", t.first.c_str(), e.what());
        }
        std::fflush(stdout);
    }
    std::printf("%d of %d tests passed\n", run - failed, run);
    honei_hip_shutdown();
    return failed ? 1 : 0;
}
