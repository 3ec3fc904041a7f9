// honei_oracle.cpp -- CPU restatement of HONEI's tags::CPU sparse-solve path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the MI355X backend:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// the library built from it.  Nothing in honei_amd/ links, loads or calls it.
//
// Every function restates one reference routine (cited file:line, paths
// relative to the HONEI source tree) with the same arithmetic order, so that
// FP results are reproducible bit-for-bit when compiled with
// -ffp-contract=off (no fused multiply-add, sequential sums as in the
// reference).  Index work (CSR -> ELL) is integer and must be bit-exact.
//
// Conventions: "f" suffix = float data, "d" = double data.  ELL matrices are
// passed as (rows, stride, threads, Aj[uint64], Ax[DT], Arl[uint64]) exactly as
// SparseMatrixELL stores them (la/sparse_matrix_ell.hh:44-129).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint64_t u64;

namespace {

template <typename DT>
struct Ell {
    u64 rows, cols, stride, ncpr, threads;
    const u64 *Aj;
    const DT *Ax;
    const u64 *Arl;
};

// ---------------------------------------------------------------------------
// ELL construction
// ---------------------------------------------------------------------------

inline u64 ceil_div_double(u64 a, u64 b) {
    // la/sparse_matrix_ell-impl.hh:89 computes ceil(double(a / double(b))).
    return (u64)std::ceil(double(a) / double(b));
}

inline u64 ell_stride(u64 rows, u64 threads) {
    // la/sparse_matrix_ell-impl.hh:97-98: 32-aligned rows*threads.
    const u64 alignment = 32;
    return alignment * (((rows * threads) + alignment - 1) / alignment);
}

inline u64 ell_slot(u64 target, u64 row, u64 threads, u64 stride) {
    // la/sparse_matrix_ell-impl.hh:113-114: slot of the target-th stored entry.
    return (target % threads) + row * threads + (target / threads) * stride;
}

// From SparseMatrix rows (la/sparse_matrix_ell-impl.hh:75-122).  Row lengths
// and the width count *stored* entries (SparseVector::used_elements), while
// only non-zero values are placed.
template <typename DT>
void ell_fill_sm(u64 rows, const u64 *row_ptr, const u64 *col, const DT *val,
                 u64 threads, u64 stride, u64 ncpr, u64 *Aj, DT *Ax, u64 *Arl) {
    std::fill(Aj, Aj + ncpr * stride, (u64)0);
    std::fill(Ax, Ax + ncpr * stride, DT(0));
    for (u64 row = 0; row < rows; ++row) {
        const u64 ue = row_ptr[row + 1] - row_ptr[row];
        Arl[row] = ceil_div_double(ue, threads);
        u64 target = 0;
        for (u64 i = row_ptr[row]; i < row_ptr[row + 1]; ++i) {
            if (val[i] != DT(0)) {
                const u64 s = ell_slot(target, row, threads, stride);
                Aj[s] = col[i];
                Ax[s] = val[i];
                ++target;
            }
        }
    }
}

// From SparseMatrixCSR (la/sparse_matrix_ell-impl.hh:124-183).  Here row
// lengths count non-zero values only; blocksize expands each CSR entry into
// `bs` consecutive columns.
template <typename DT>
void ell_shape_csr(u64 rows, const u64 *Ar, const DT *Ax, u64 bs, u64 threads,
                   u64 *stride, u64 *ncpr) {
    u64 width = 1;
    for (u64 r = 0; r < rows; ++r) {
        u64 ue = 0;
        for (u64 j = Ar[r] * bs; j < Ar[r + 1] * bs; ++j)
            if (Ax[j] != DT(0)) ++ue;
        if (ue > width) width = ue;
    }
    *ncpr = ceil_div_double(width, threads);
    *stride = ell_stride(rows, threads);
}

template <typename DT>
void ell_fill_csr(u64 rows, const u64 *Ar, const u64 *Ajc, const DT *Axc, u64 bs,
                  u64 threads, u64 stride, u64 ncpr, u64 *Aj, DT *Ax, u64 *Arl) {
    std::fill(Aj, Aj + ncpr * stride, (u64)0);
    std::fill(Ax, Ax + ncpr * stride, DT(0));
    for (u64 row = 0; row < rows; ++row) {
        u64 ue = 0;
        for (u64 j = Ar[row] * bs; j < Ar[row + 1] * bs; ++j)
            if (Axc[j] != DT(0)) ++ue;
        Arl[row] = ceil_div_double(ue, threads);
        u64 target = 0;
        for (u64 i = Ar[row]; i < Ar[row + 1]; ++i) {
            for (u64 b = 0; b < bs; ++b) {
                const DT v = Axc[i * bs + b];
                if (v != DT(0)) {
                    const u64 s = ell_slot(target, row, threads, stride);
                    Aj[s] = Ajc[i] + b;
                    Ax[s] = v;
                    ++target;
                }
            }
        }
    }
}

// Row lengths of an ELL matrix built from raw arrays
// (la/sparse_matrix_ell-impl.hh:186-208): count chunks whose thread-0 value
// is non-zero, stopping at the first chunk holding any zero.
template <typename DT>
void ell_row_length(u64 rows, u64 stride, u64 threads, u64 size, const DT *Ax, u64 *Arl) {
    for (u64 row = 0; row < rows; ++row) {
        u64 count = 0;
        bool stop = false;
        for (u64 i = row * threads; i < size && !stop; i += stride) {
            for (u64 t = 0; t < threads; ++t) {
                if (Ax[i + t] == DT(0)) { stop = true; break; }
                if (t == 0) ++count;
            }
        }
        Arl[row] = count;
    }
}

// ---------------------------------------------------------------------------
// Products and BLAS-1 (all sequential, reference order)
// ---------------------------------------------------------------------------

// la/product.hh:1450-1482 (Product<tags::CPU::Generic>, ELL), same order as
// la/product.hh:148-206: k-major over Arl[row] chunks, then threads.
template <typename DT>
void product_ell(DT *y, const DT *x, const Ell<DT> &A, u64 row_start, u64 row_end) {
    const u64 T = A.threads, stride = A.stride;
    for (u64 row = row_start; row < row_end; ++row) {
        DT sum(0);
        const u64 *aj = A.Aj + row * T;
        const DT *ax = A.Ax + row * T;
        const u64 max = A.Arl[row];
        for (u64 n = 0; n < max; ++n) {
            for (u64 t = 0; t < T; ++t) sum += ax[t] * x[aj[t]];
            aj += stride;
            ax += stride;
        }
        y[row] = sum;
    }
}

// math/defect.hh:449-504 (Defect<tags::CPU::Generic>, ELL): r = rhs - A x.
template <typename DT>
void defect_ell(DT *r, const DT *rhs, const DT *x, const Ell<DT> &A) {
    const u64 T = A.threads, stride = A.stride;
    for (u64 row = 0; row < A.rows; ++row) {
        DT sum(0);
        const u64 *aj = A.Aj + row * T;
        const DT *ax = A.Ax + row * T;
        const u64 max = A.Arl[row];
        for (u64 n = 0; n < max; ++n) {
            for (u64 t = 0; t < T; ++t) sum += ax[t] * x[aj[t]];
            aj += stride;
            ax += stride;
        }
        r[row] = rhs[row] - sum;
    }
}

// la/dot_product.hh:81-84: sequential sum.
template <typename DT>
DT dot(const DT *x, const DT *y, u64 n) {
    DT r(0);
    for (u64 i = 0; i < n; ++i) r += x[i] * y[i];
    return r;
}

// la/norm.hh:290-310: Norm<l2,false> = DotProduct(x,x); root variant = sqrt.
template <typename DT>
DT norm2(const DT *x, u64 n) { return dot(x, x, n); }

// la/scaled_sum.hh:73-88: x += b*y.
template <typename DT>
void scaled_sum2(DT *x, const DT *y, DT b, u64 n) {
    for (u64 i = 0; i < n; ++i) x[i] += y[i] * b;
}
// la/scaled_sum.hh:91-110: r = x + b*y.
template <typename DT>
void scaled_sum3(DT *r, const DT *x, const DT *y, DT b, u64 n) {
    for (u64 i = 0; i < n; ++i) r[i] = x[i] + y[i] * b;
}
// la/scaled_sum.hh:177-196: a += b .* c.
template <typename DT>
void scaled_sum_v(DT *a, const DT *b, const DT *c, u64 n) {
    for (u64 i = 0; i < n; ++i) a[i] += b[i] * c[i];
}
template <typename DT> void vsum(DT *x, const DT *y, u64 n) { for (u64 i = 0; i < n; ++i) x[i] += y[i]; }
template <typename DT> void vdiff(DT *x, const DT *y, u64 n) { for (u64 i = 0; i < n; ++i) x[i] -= y[i]; }
template <typename DT> void vscale(DT *x, DT a, u64 n) { for (u64 i = 0; i < n; ++i) x[i] *= a; }
template <typename DT> void vprod(DT *r, const DT *x, const DT *y, u64 n) { for (u64 i = 0; i < n; ++i) r[i] = x[i] * y[i]; }
template <typename DT> void vfill(DT *x, DT v, u64 n) { for (u64 i = 0; i < n; ++i) x[i] = v; }
template <typename DT> void vcopy(DT *d, const DT *s, u64 n) { std::memcpy(d, s, n * sizeof(DT)); }

// ---------------------------------------------------------------------------
// Q1 Poisson operator (math/fill_matrix.hh:36-133) and band -> SparseMatrix
// (la/sparse_matrix.hh:221-300)
// ---------------------------------------------------------------------------

enum { LL = 0, LD, LU, DL, DD, DU, UL, UD, UU };

template <typename DT>
void q1_bands(u64 N, DT *bands, int bc) {
    DT *b[9];
    for (int k = 0; k < 9; ++k) b[k] = bands + (u64)k * N;
    const u64 M = (u64)std::sqrt((double)N);
    for (u64 i = 1; i <= N; ++i) {
        bool dirichlet, neumann = false;
        if (bc == 0) {
            dirichlet = (i <= M || i > N - M || (i % M) == 1 || (i % M) == 0);
        } else {
            dirichlet = (i <= M || i > N - M || (i % M) == 1);
            neumann = !dirichlet && (i % M) == 0 && i > M && i < N;
        }
        const u64 j = i - 1;
        if (dirichlet) {
            for (int k = 0; k < 9; ++k) b[k][j] = DT(0);
            b[DD][j] = DT(1.0);
        } else if (neumann) {
            b[DD][j] = DT(4.0 / 3.0);
            b[LL][j] = DT(-1.0 / 3.0);
            b[LD][j] = DT(-1.0 / 6.0);
            b[DL][j] = DT(-1.0 / 3.0);
            b[UL][j] = DT(-1.0 / 3.0);
            b[UD][j] = DT(-1.0 / 6.0);
            b[LU][j] = b[DU][j] = b[UU][j] = DT(0.0);
        } else {
            b[DD][j] = DT(8.0 / 3.0);
            b[LL][j] = b[LD][j] = b[LU][j] = DT(-1.0 / 3.0);
            b[DL][j] = b[DU][j] = DT(-1.0 / 3.0);
            b[UL][j] = b[UD][j] = b[UU][j] = DT(-1.0 / 3.0);
        }
    }
}

// Product<tags::CPU>::value(BandedMatrixQx<Q1Type>, DenseVector)
// (la/product.hh:371-400): DD first, then LL, LD, LU, DL, DU, UL, UD, UU,
// each term guarded by its column range; Defect<tags::CPU> on the banded
// matrix is rhs - product (math/defect.hh:194-199).
template <typename DT>
void banded_q1(u64 N, const DT *bands, const DT *rhs, const DT *x, DT *y) {
    const DT *b[9];
    for (int k = 0; k < 9; ++k) b[k] = bands + (u64)k * N;
    const long n = (long)N, root = (long)std::sqrt((double)N);
    for (long i = 0; i < n; ++i) {
        DT r = b[DD][i] * x[i];
        if (i - root - 1 >= 0) r += b[LL][i] * x[i - root - 1];
        if (i - root >= 0) r += b[LD][i] * x[i - root];
        if (i - root + 1 >= 0) r += b[LU][i] * x[i - root + 1];
        if (i - 1 >= 0) r += b[DL][i] * x[i - 1];
        if (i + 1 < n) r += b[DU][i] * x[i + 1];
        if (i + root - 1 < n) r += b[UL][i] * x[i + root - 1];
        if (i + root < n) r += b[UD][i] * x[i + root];
        if (i + root + 1 < n) r += b[UU][i] * x[i + root + 1];
        y[i] = rhs ? rhs[i] - r : r;
    }
}

// Variable-coefficient Q1 operator (no reference counterpart; the device
// assembly in honei_amd/csrc/q1_kernels.hip follows the same formulas):
// kappa per cell ((side-1)^2, row-major), bilinear element stiffness
// kappa/6 * [[4,-1,-2,-1],[-1,4,-1,-2],[-2,-1,4,-1],[-1,-2,-1,4]] summed over
// the four cells around node (i, j) -- cells sw=(i-1,j-1), se=(i-1,j),
// nw=(i,j-1), ne=(i,j); Dirichlet rows identity.  kappa = 1 gives the
// constant stencil of q1_bands (8/3, -1/3).
template <typename DT>
void q1_coeff_bands(u64 side, const double *kappa, DT *bands) {
    const u64 N = side * side, nc = side - 1;
    DT *b[9];
    for (int k = 0; k < 9; ++k) b[k] = bands + (u64)k * N;
    for (u64 r = 0; r < N; ++r) {
        const u64 i = r / side, j = r % side;
        for (int k = 0; k < 9; ++k) b[k][r] = DT(0);
        if (i == 0 || i == side - 1 || j == 0 || j == side - 1) {
            b[DD][r] = DT(1.0);
            continue;
        }
        const double sw = kappa[(i - 1) * nc + (j - 1)], se = kappa[(i - 1) * nc + j];
        const double nw = kappa[i * nc + (j - 1)], ne = kappa[i * nc + j];
        b[LL][r] = (DT)(-sw / 3.0);
        b[LD][r] = (DT)(-(sw + se) / 6.0);
        b[LU][r] = (DT)(-se / 3.0);
        b[DL][r] = (DT)(-(sw + nw) / 6.0);
        b[DD][r] = (DT)(2.0 * (sw + se + nw + ne) / 3.0);
        b[DU][r] = (DT)(-(se + ne) / 6.0);
        b[UL][r] = (DT)(-nw / 3.0);
        b[UD][r] = (DT)(-(nw + ne) / 6.0);
        b[UU][r] = (DT)(-ne / 3.0);
    }
}

// Coarse cell (I, J) from fine cells (2I..2I+1, 2J..2J+1): averaging 0
// arithmetic, 1 geometric, 2 harmonic mean (rediscretised hierarchies).
void coarsen_coeff(const double *fine, u64 nf, double *coarse, u64 nc, int averaging) {
    for (u64 I = 0; I < nc; ++I)
        for (u64 J = 0; J < nc; ++J) {
            const double a = fine[(2 * I) * nf + 2 * J], b = fine[(2 * I) * nf + 2 * J + 1];
            const double c = fine[(2 * I + 1) * nf + 2 * J], d = fine[(2 * I + 1) * nf + 2 * J + 1];
            double m;
            if (averaging == 1) m = std::exp(0.25 * (((std::log(a) + std::log(b)) + std::log(c)) + std::log(d)));
            else if (averaging == 2) m = 4.0 / (((1.0 / a + 1.0 / b) + 1.0 / c) + 1.0 / d);
            else m = 0.25 * (((a + b) + c) + d);
            coarse[I * nc + J] = m;
        }
}

// Each band k contributes entry (i, i + off_k) for i in [lo_k, hi_k), only when
// non-zero; rows come out sorted by column because bands are visited in
// ascending offset order.
template <typename DT>
u64 banded_q1_to_csr(u64 N, const DT *bands, u64 *row_ptr, u64 *col, DT *val) {
    const long root = (long)std::sqrt((double)N);
    const long n = (long)N;
    const long off[9] = {-root - 1, -root, -root + 1, -1, 0, 1, root - 1, root, root + 1};
    const long lo[9] = {root + 1, root, root - 1, 1, 0, 0, 0, 0, 0};
    const long hi[9] = {n, n, n, n, n, n - 1, n - root + 1, n - root, n - root - 1};
    u64 nnz = 0;
    for (long i = 0; i < n; ++i) {
        if (row_ptr) row_ptr[i] = nnz;
        for (int k = 0; k < 9; ++k) {
            if (i < lo[k] || i >= hi[k]) continue;
            const DT v = bands[(u64)k * N + i];
            if (v == DT(0)) continue;
            if (col) { col[nnz] = (u64)(i + off[k]); val[nnz] = v; }
            ++nnz;
        }
    }
    if (row_ptr) row_ptr[N] = nnz;
    return nnz;
}

// ---------------------------------------------------------------------------
// Geometric transfers (math/restriction.hh:40-172, math/prolongation.hh:41-139,
// math/apply_dirichlet_boundaries.hh:28-85)
// ---------------------------------------------------------------------------

template <typename DT>
void apply_dirichlet(DT *g, u64 N, const u64 *mask) {
    const u64 M = (u64)(int)std::sqrt((double)N);
    if (mask[0] == 2) g[0] = 0;
    if (mask[1] == 2) g[M - 1] = 0;
    if (mask[2] == 2) g[N - 1] = 0;
    if (mask[3] == 2) g[N - M] = 0;
    if (mask[4] == 2) for (u64 i = 1; i < M - 1; ++i) g[i] = 0;
    if (mask[5] == 2) for (u64 i = 2 * M - 1; i < N - M; i += M) g[i] = 0;
    if (mask[6] == 2) for (u64 i = N - M + 1; i < N - 1; ++i) g[i] = 0;
    if (mask[7] == 2) for (u64 i = M; i < N - M; i += M) g[i] = 0;
}

// Full weighting (transpose of bilinear interpolation) accumulated in the
// reference's loop order.  Grid coordinates: fine node (r, c) has index
// r*nf + c, coarse node (r, c) index r*nc + c.
template <typename DT>
void restriction(DT *coarse, u64 n_coarse, const DT *fine, u64 n_fine, const u64 *mask) {
    (void)n_coarse;
    const u64 nxf = (u64)std::sqrt((DT)n_fine) - 1;
    const u64 m = nxf / 2;        // coarse cells per side
    const u64 nf = nxf + 1, nc = m + 1;
    auto F = [&](u64 r, u64 c) -> DT { return fine[r * nf + c]; };
    auto C = [&](u64 r, u64 c) -> DT & { return coarse[r * nc + c]; };
    for (u64 r = 0; r < nc; ++r)
        for (u64 c = 0; c < nc; ++c) C(r, c) = F(2 * r, 2 * c);
    for (u64 i = 1; i <= m; ++i) {
        // fine row 2i-1 (between coarse rows i-1 and i), even columns
        for (u64 j = 0; j < nc; ++j) C(i, j) += DT(0.5) * F(2 * i - 1, 2 * j);
        for (u64 j = 0; j < nc; ++j) C(i - 1, j) += DT(0.5) * F(2 * i - 1, 2 * j);
        // fine column 2i-1 (between coarse columns i-1 and i), even rows
        for (u64 j = 0; j < nc; ++j) C(j, i - 1) += DT(0.5) * F(2 * j, 2 * i - 1);
        for (u64 j = 0; j < nc; ++j) C(j, i) += DT(0.5) * F(2 * j, 2 * i - 1);
        // cell centres of fine row 2i-1
        for (u64 j = 0; j < m; ++j) C(i - 1, j) += DT(0.25) * F(2 * i - 1, 2 * j + 1);
        for (u64 j = 0; j < m; ++j) C(i - 1, j + 1) += DT(0.25) * F(2 * i - 1, 2 * j + 1);
        for (u64 j = 0; j < m; ++j) C(i, j) += DT(0.25) * F(2 * i - 1, 2 * j + 1);
        for (u64 j = 0; j < m; ++j) C(i, j + 1) += DT(0.25) * F(2 * i - 1, 2 * j + 1);
    }
    apply_dirichlet(coarse, nc * nc, mask);
}

template <typename DT>
void prolongation(DT *fine, u64 n_fine, const DT *coarse, u64 n_coarse, const u64 *mask) {
    (void)n_coarse;
    const u64 nxf = (u64)std::sqrt((DT)n_fine) - 1;
    const u64 m = nxf / 2;
    const u64 nf = nxf + 1, nc = m + 1;
    auto F = [&](u64 r, u64 c) -> DT & { return fine[r * nf + c]; };
    auto C = [&](u64 r, u64 c) -> DT { return coarse[r * nc + c]; };
    for (u64 i = 0; i < n_fine; ++i) fine[i] = DT(0);
    for (u64 r = 0; r < nc; ++r)
        for (u64 c = 0; c < nc; ++c) F(2 * r, 2 * c) = C(r, c);
    for (u64 i = 1; i <= m; ++i) {
        for (u64 j = 0; j < nc; ++j) {
            F(2 * j, 2 * i - 1) += C(j, i - 1) * DT(0.5);
            F(2 * j, 2 * i - 1) += C(j, i) * DT(0.5);
        }
        for (u64 j = 0; j < nc; ++j) {
            F(2 * i - 1, 2 * j) += F(2 * i - 2, 2 * j) * DT(0.5);
            F(2 * i - 1, 2 * j) += F(2 * i, 2 * j) * DT(0.5);
        }
    }
    for (u64 i = 1; i <= m; ++i) {
        for (u64 j = 0; j < m; ++j) {
            F(2 * i - 1, 2 * j + 1) += F(2 * i - 2, 2 * j) * DT(0.25);
            F(2 * i - 1, 2 * j + 1) += F(2 * i - 2, 2 * j + 2) * DT(0.25);
            F(2 * i - 1, 2 * j + 1) += F(2 * i, 2 * j) * DT(0.25);
            F(2 * i - 1, 2 * j + 1) += F(2 * i, 2 * j + 2) * DT(0.25);
        }
    }
    apply_dirichlet(fine, n_fine, mask);
}

// math/prolongation_matrix.hh:28-100 with natural ordering
// (math/reordering.hh:36-48): P is n_f x n_c, weights 1, 1/2, 1/4.  Entries of
// a row are emitted sorted by coarse column (SparseVector keeps indices
// sorted).
u64 prolmat_csr(u64 fine_dim, u64 *row_ptr, u64 *col, double *val) {
    const u64 cd = (fine_dim + 1) / 2;
    auto cidx = [&](u64 r, u64 c) { return r * cd + c; };
    u64 nnz = 0;
    for (u64 i = 0; i < fine_dim; ++i) {
        for (u64 j = 0; j < fine_dim; ++j) {
            const u64 row = i * fine_dim + j;
            if (row_ptr) row_ptr[row] = nnz;
            u64 cs[4];
            double ws[4];
            int k = 0;
            if (i % 2 == 0 && j % 2 == 0) {
                cs[k] = cidx(i / 2, j / 2); ws[k++] = 1.0;
            } else if (i % 2 == 0) {
                cs[k] = cidx(i / 2, (j - 1) / 2); ws[k++] = 0.5;
                cs[k] = cidx(i / 2, (j + 1) / 2); ws[k++] = 0.5;
            } else if (j % 2 == 0) {
                cs[k] = cidx((i - 1) / 2, j / 2); ws[k++] = 0.5;
                cs[k] = cidx((i + 1) / 2, j / 2); ws[k++] = 0.5;
            } else {
                cs[k] = cidx((i - 1) / 2, (j - 1) / 2); ws[k++] = 0.25;
                cs[k] = cidx((i - 1) / 2, (j + 1) / 2); ws[k++] = 0.25;
                cs[k] = cidx((i + 1) / 2, (j - 1) / 2); ws[k++] = 0.25;
                cs[k] = cidx((i + 1) / 2, (j + 1) / 2); ws[k++] = 0.25;
            }
            for (int q = 0; q < k; ++q) {
                if (col) { col[nnz] = cs[q]; val[nnz] = ws[q]; }
                ++nnz;
            }
        }
    }
    if (row_ptr) row_ptr[fine_dim * fine_dim] = nnz;
    return nnz;
}

// ---------------------------------------------------------------------------
// Solvers
// ---------------------------------------------------------------------------

// math/cg.hh:63-133 (CGSolver<Tag, NONE>).
template <typename DT>
void cg_solver(const Ell<DT> &A, const DT *b, DT *x, u64 max_iters, u64 *used_iters,
               DT eps_relative, DT *history) {
    const u64 n = A.rows;
    std::vector<DT> p(n), r(n), v(n);
    const DT eps = std::numeric_limits<DT>::epsilon();
    defect_ell(r.data(), b, x, A);
    vcopy(p.data(), r.data(), n);
    DT alpha = norm2(r.data(), n);
    const DT initial_defect = std::sqrt(alpha);
    u64 it = 0;
    *used_iters = 0;
    while (it < max_iters) {
        product_ell(v.data(), p.data(), A, 0, n);
        const DT temp = dot(v.data(), p.data(), n);
        const DT lambda = alpha / (std::abs(temp) > eps ? temp : eps);
        scaled_sum2(x, p.data(), lambda, n);
        const DT mlambda = -lambda;
        scaled_sum2(r.data(), v.data(), mlambda, n);
        const DT alpha_old = alpha;
        alpha = norm2(r.data(), n);
        const DT talpha = alpha / (std::abs(alpha_old) > eps ? alpha_old : eps);
        vscale(p.data(), talpha, n);
        vsum(p.data(), r.data(), n);
        ++it;
        const DT current = std::sqrt(alpha);
        if (history) history[it - 1] = current;
        if (current < eps_relative * initial_defect) { *used_iters = it; break; }
        if (current < eps_relative) { *used_iters = it; break; }
        if (it == max_iters) *used_iters = it;
    }
}

// ConjugateGradients<Tag, JAC>::value on ELL (math/conjugate_gradients.hh:
// 1110-1154) with its cg_kernel (math/conjugate_gradients.hh:884-913):
// Jacobi-preconditioned CG, z = dinv .* r.  Stops once ||r|| < initial *
// eps_relative; used_iters is written only when the loop converges or
// reaches max_iters (it stays untouched for max_iters = 0).  history (if
// given) receives ||r|| after every iteration.
template <typename DT>
void pcg_jacobi(const Ell<DT> &A, const DT *b, DT *x, const DT *dinv, u64 max_iters, u64 *used_iters,
                DT eps_relative, DT *history) {
    const u64 n = A.rows;
    const DT eps = std::numeric_limits<DT>::epsilon();
    std::vector<DT> r(n), z(n), d(n), t0(n);
    defect_ell(r.data(), b, x, A);
    const DT initial = std::sqrt(norm2(r.data(), n));
    vprod(z.data(), dinv, r.data(), n);
    vcopy(d.data(), z.data(), n);
    for (u64 i = 0; i < max_iters; ++i) {
        DT beta = dot(r.data(), z.data(), n);
        product_ell(t0.data(), d.data(), A, 0, n);
        DT alpha = dot(d.data(), t0.data(), n);
        alpha = std::abs(alpha) >= eps ? beta / alpha : beta / eps;
        scaled_sum2(x, d.data(), alpha, n);
        scaled_sum2(r.data(), t0.data(), DT(-alpha), n);
        vprod(z.data(), dinv, r.data(), n);
        const DT zr = dot(z.data(), r.data(), n);
        beta = std::abs(beta) >= eps ? zr / beta : zr / eps;
        vcopy(t0.data(), d.data(), n);
        scaled_sum3(d.data(), z.data(), t0.data(), beta, n);
        const DT current = std::sqrt(norm2(r.data(), n));
        if (history) history[i] = current;
        if (current < initial * eps_relative) { *used_iters = i + 1; break; }
        if (i == max_iters - 1) *used_iters = i + 1;
    }
}

// math/conjugate_gradients.hh:154-201 (cg_kernel, ELL, gradient form) driven
// by the konv_rad loop of math/conjugate_gradients.hh:587-623: x starts at 0
// and iterations stop once ||x||^2 grows by no more than konv_rad.
template <typename DT>
u64 cg_konv(const Ell<DT> &A, const DT *rhs, DT *x, double konv_rad) {
    const u64 n = A.rows;
    const DT eps = std::numeric_limits<DT>::epsilon();
    vfill(x, DT(0), n);
    std::vector<DT> g(n), u(n), energy(n), uc(n);
    product_ell(g.data(), x, A, 0, n);
    vdiff(g.data(), rhs, n);
    vcopy(u.data(), g.data(), n);
    vscale(u.data(), DT(-1.), n);
    DT norm_x = norm2(x, n);
    DT norm_x_last = DT(norm_x - 1);
    u64 iters = 0;
    while (norm_x - norm_x_last > konv_rad) {
        DT upper = norm2(g.data(), n);
        product_ell(energy.data(), u.data(), A, 0, n);
        DT lower = dot(energy.data(), u.data(), n);
        vcopy(uc.data(), u.data(), n);
        if (std::fabs(lower) >= eps) vscale(uc.data(), DT(upper / lower), n);
        else vscale(uc.data(), DT(upper / eps), n);
        vcopy(energy.data(), uc.data(), n);
        vsum(energy.data(), x, n);
        lower = upper;
        product_ell(g.data(), energy.data(), A, 0, n);
        vdiff(g.data(), rhs, n);
        upper = norm2(g.data(), n);
        if (std::fabs(lower) >= eps) vscale(u.data(), DT(upper / lower), n);
        else vscale(u.data(), DT(upper / eps), n);
        vdiff(u.data(), g.data(), n);
        vcopy(x, energy.data(), n);
        ++iters;
        norm_x_last = norm_x;
        norm_x = norm2(x, n);
    }
    return iters;
}

}  // namespace

// ---------------------------------------------------------------------------
// Multigrid: math/multigrid.hh:1936-2250 (_multigrid_kernel, JAC, V-cycle),
// FIXED driver math/multigrid.hh:828-870, MIXED driver
// math/multigrid.hh:2395-2536.
// ---------------------------------------------------------------------------

extern "C" {

typedef struct {
    u64 rows, cols, stride, ncpr, threads;
    const u64 *Aj;
    const void *Ax;
    const u64 *Arl;
} oracle_ell;

typedef struct {
    oracle_ell A;       // system matrix on this level
    const void *dinv;   // omega / diag(A)
    oracle_ell prol;    // PROLMAT: P from level-1 into this level
    oracle_ell res;     // PROLMAT: R = P^T from this level into level-1
} oracle_mg_level;

typedef struct {
    u64 min_level, max_level, n_max_iter, n_pre_smooth, n_post_smooth;
    int initial_zero, convergence_check, is_smoother, prolmat;
    double tolerance, adapt_correction_factor;
    u64 mask[8];
    // Outer loop of the MIXED driver: zero the inner start vector on every
    // outer step instead of carrying the previous inner solution over
    // (reference behaviour is to carry it over).
    int zero_inner_start;
    u64 max_outer;      // 0 = unbounded (reference)
} oracle_mg_config;

}  // extern "C"

namespace {

template <typename DT>
Ell<DT> as_ell(const oracle_ell &e) {
    Ell<DT> r{e.rows, e.cols, e.stride, e.ncpr, e.threads, e.Aj, (const DT *)e.Ax, e.Arl};
    return r;
}

template <typename DT>
struct MGState {
    std::vector<std::vector<DT>> c, d, rhs, x;
};

// Jacobi<Tag>::value first sweep from zero (math/jacobi.hh:481-493):
// x = dinv .* rhs.  Further sweeps (math/jacobi.hh:268-279, 495-513):
// t = rhs - A x; x += t .* dinv.
template <typename DT>
void jacobi_first(DT *x, const DT *dinv, const DT *rhs, u64 n) { vprod(x, dinv, rhs, n); }

template <typename DT>
void jacobi_sweeps(DT *x, const Ell<DT> &A, const DT *rhs, const DT *dinv, u64 iters) {
    std::vector<DT> t(A.rows);
    for (u64 i = 0; i < iters; ++i) {
        defect_ell(t.data(), rhs, x, A);
        scaled_sum_v(x, t.data(), dinv, A.rows);
    }
}

template <typename DT>
void restrict_level(const oracle_mg_level *L, const oracle_mg_config &cfg, u64 fine_level,
                    DT *coarse, u64 nc, const DT *fine, u64 nf) {
    if (cfg.prolmat) {
        Ell<DT> R = as_ell<DT>(L[fine_level].res);
        product_ell(coarse, fine, R, 0, R.rows);
    } else {
        restriction(coarse, nc, fine, nf, cfg.mask);
    }
}

template <typename DT>
void prolongate_level(const oracle_mg_level *L, const oracle_mg_config &cfg, u64 fine_level,
                      DT *fine, u64 nf, const DT *coarse, u64 nc) {
    if (cfg.prolmat) {
        Ell<DT> P = as_ell<DT>(L[fine_level].prol);
        product_ell(fine, coarse, P, 0, P.rows);
    } else {
        prolongation(fine, nf, coarse, nc, cfg.mask);
    }
}

// One call of _multigrid_kernel.  The state vectors are owned by the caller
// so the MIXED driver can keep x[max] between calls as the reference does.
template <typename DT>
u64 mg_kernel(const oracle_mg_level *L, const oracle_mg_config &cfg, MGState<DT> &S,
              double *defect_out) {
    const u64 maxl = cfg.max_level, minl = cfg.min_level;
    auto n_of = [&](u64 l) { return L[l].A.rows; };
    Ell<DT> Amax = as_ell<DT>(L[maxl].A);
    if (cfg.initial_zero) {
        S.d[maxl] = S.rhs[maxl];
    } else {
        defect_ell(S.d[maxl].data(), S.rhs[maxl].data(), S.x[maxl].data(), Amax);
    }
    DT defect, initial_defect;
    if (!cfg.is_smoother) {
        defect = std::sqrt(norm2(S.d[maxl].data(), n_of(maxl)));
        initial_defect = defect;
    } else {
        defect = DT(1e8);
        initial_defect = defect;
    }
    u64 inner_iterations = 0;
    if (cfg.convergence_check && defect <= DT(cfg.tolerance)) {
        // nothing to do
    } else {
        u64 iter;
        for (iter = 1; iter <= cfg.n_max_iter; ++iter) {
            u64 level = maxl;
            if (minl != maxl) {
                // restriction loop
                bool started = true;
                while (true) {
                    Ell<DT> A = as_ell<DT>(L[level].A);
                    const DT *dinv = (const DT *)L[level].dinv;
                    const u64 n = n_of(level);
                    if (started) {
                        jacobi_first(S.c[level].data(), dinv, S.d[level].data(), n);
                        jacobi_sweeps(S.c[level].data(), A, S.d[level].data(), dinv, cfg.n_pre_smooth - 1);
                        vsum(S.x[level].data(), S.c[level].data(), n);
                    } else {
                        jacobi_first(S.x[level].data(), dinv, S.d[level].data(), n);
                        jacobi_sweeps(S.x[level].data(), A, S.d[level].data(), dinv, cfg.n_pre_smooth - 1);
                    }
                    defect_ell(S.d[level].data(), S.rhs[level].data(), S.x[level].data(), A);
                    --level;
                    restrict_level(L, cfg, level + 1, S.d[level].data(), n_of(level),
                                   S.d[level + 1].data(), n_of(level + 1));
                    S.rhs[level] = S.d[level];
                    if (level == minl) break;
                    started = false;
                }
            }
            // coarse grid correction: ConjugateGradients konv_rad = FLT/DBL eps
            {
                Ell<DT> A = as_ell<DT>(L[level].A);
                cg_konv(A, S.d[level].data(), S.x[level].data(),
                        (double)std::numeric_limits<DT>::epsilon());
                if (minl == maxl) {
                    defect_ell(S.d[level].data(), S.rhs[level].data(), S.x[level].data(), A);
                    break;
                }
            }
            // prolongation loop
            while (true) {
                ++level;
                const u64 n = n_of(level);
                Ell<DT> A = as_ell<DT>(L[level].A);
                const DT *dinv = (const DT *)L[level].dinv;
                prolongate_level(L, cfg, level, S.c[level].data(), n, S.x[level - 1].data(), n_of(level - 1));
                DT alpha;
                if (cfg.adapt_correction_factor == 0.0) {
                    const DT d1 = dot(S.d[level].data(), S.c[level].data(), n);
                    std::vector<DT> t(n, DT(0));
                    product_ell(t.data(), S.c[level].data(), A, 0, n);
                    const DT d2 = dot(t.data(), S.c[level].data(), n);
                    alpha = DT(d1 / d2);
                } else {
                    alpha = DT(cfg.adapt_correction_factor);
                }
                scaled_sum2(S.x[level].data(), S.c[level].data(), alpha, n);
                // post-smoothing uses n_pre_smooth sweeps (math/multigrid.hh:2168)
                jacobi_sweeps(S.x[level].data(), A, S.rhs[level].data(), dinv, cfg.n_pre_smooth);
                defect_ell(S.d[level].data(), S.rhs[level].data(), S.x[level].data(), A);
                if (level == maxl) break;
            }
            if (!cfg.is_smoother) {
                defect = std::sqrt(norm2(S.d[maxl].data(), n_of(maxl)));
                if (defect <= initial_defect * DT(cfg.tolerance)) break;
            }
        }
        inner_iterations = (iter > cfg.n_max_iter) ? iter - 1 : iter;
    }
    if (defect_out) *defect_out = (double)defect;
    return inner_iterations;
}

template <typename DT>
void mg_alloc(const oracle_mg_level *L, const oracle_mg_config &cfg, MGState<DT> &S) {
    const u64 nl = cfg.max_level + 1;
    S.c.assign(nl, {}); S.d.assign(nl, {}); S.rhs.assign(nl, {}); S.x.assign(nl, {});
    for (u64 l = cfg.min_level; l <= cfg.max_level; ++l) {
        const u64 n = L[l].A.rows;
        S.c[l].assign(n, DT(0)); S.d[l].assign(n, DT(0));
        S.rhs[l].assign(n, DT(0)); S.x[l].assign(n, DT(0));
    }
}

// Boundary start values of the MG drivers (math/multigrid.hh:2423-2440):
// first grid row, last grid row and first grid column take b; the last
// column is left as given.
template <typename DT>
void mg_boundary_guess(DT *x, const DT *b, u64 n) {
    const u64 s = (u64)std::sqrt((double)n);
    for (u64 i = 0; i < s; ++i) x[i] = b[i];
    for (u64 i = n - s; i < n; ++i) x[i] = b[i];
    for (u64 i = 0; i < n; i += s) x[i] = b[i];
}

}  // namespace

extern "C" {

// ---- ELL construction ------------------------------------------------------
void oracle_ell_shape_sm(u64 rows, const u64 *row_ptr, u64 threads, u64 *stride, u64 *ncpr) {
    u64 width = 1;
    for (u64 r = 0; r < rows; ++r) width = std::max(width, row_ptr[r + 1] - row_ptr[r]);
    *ncpr = ceil_div_double(width, threads);
    *stride = ell_stride(rows, threads);
}
void oracle_ell_fill_sm_f(u64 rows, const u64 *rp, const u64 *col, const float *val, u64 T, u64 stride, u64 ncpr, u64 *Aj, float *Ax, u64 *Arl) { ell_fill_sm(rows, rp, col, val, T, stride, ncpr, Aj, Ax, Arl); }
void oracle_ell_fill_sm_d(u64 rows, const u64 *rp, const u64 *col, const double *val, u64 T, u64 stride, u64 ncpr, u64 *Aj, double *Ax, u64 *Arl) { ell_fill_sm(rows, rp, col, val, T, stride, ncpr, Aj, Ax, Arl); }
void oracle_ell_shape_csr_f(u64 rows, const u64 *Ar, const float *Ax, u64 bs, u64 T, u64 *stride, u64 *ncpr) { ell_shape_csr(rows, Ar, Ax, bs, T, stride, ncpr); }
void oracle_ell_shape_csr_d(u64 rows, const u64 *Ar, const double *Ax, u64 bs, u64 T, u64 *stride, u64 *ncpr) { ell_shape_csr(rows, Ar, Ax, bs, T, stride, ncpr); }
void oracle_ell_fill_csr_f(u64 rows, const u64 *Ar, const u64 *Ajc, const float *Axc, u64 bs, u64 T, u64 stride, u64 ncpr, u64 *Aj, float *Ax, u64 *Arl) { ell_fill_csr(rows, Ar, Ajc, Axc, bs, T, stride, ncpr, Aj, Ax, Arl); }
void oracle_ell_fill_csr_d(u64 rows, const u64 *Ar, const u64 *Ajc, const double *Axc, u64 bs, u64 T, u64 stride, u64 ncpr, u64 *Aj, double *Ax, u64 *Arl) { ell_fill_csr(rows, Ar, Ajc, Axc, bs, T, stride, ncpr, Aj, Ax, Arl); }
void oracle_ell_row_length_f(u64 rows, u64 stride, u64 T, u64 size, const float *Ax, u64 *Arl) { ell_row_length(rows, stride, T, size, Ax, Arl); }
void oracle_ell_row_length_d(u64 rows, u64 stride, u64 T, u64 size, const double *Ax, u64 *Arl) { ell_row_length(rows, stride, T, size, Ax, Arl); }

// ---- products ----------------------------------------------------------------
#define ORACLE_ELL_ARGS(DT) u64 rows, u64 stride, u64 T, const u64 *Aj, const DT *Ax, const u64 *Arl
#define ORACLE_MAKE_ELL(DT) Ell<DT> A{rows, rows, stride, 0, T, Aj, Ax, Arl}

void oracle_product_ell_f(float *y, const float *x, ORACLE_ELL_ARGS(float), u64 rs, u64 re) { ORACLE_MAKE_ELL(float); product_ell(y, x, A, rs, re ? re : rows); }
void oracle_product_ell_d(double *y, const double *x, ORACLE_ELL_ARGS(double), u64 rs, u64 re) { ORACLE_MAKE_ELL(double); product_ell(y, x, A, rs, re ? re : rows); }

}  // extern "C"

// mc::Product (la/product.hh:2114-2150): `parts` contiguous row blocks of
// rows/parts rows each (last block takes the remainder), one per thread.
template <typename DT>
static void product_ell_mc(DT *y, const DT *x, const Ell<DT> &A, int parts) {
    std::vector<u64> lim(parts + 1);
    lim[0] = 0;
    for (int i = 1; i < parts; ++i) lim[i] = lim[i - 1] + A.rows / parts;
    lim[parts] = A.rows;
#pragma omp parallel for schedule(static, 1) num_threads(parts)
    for (int i = 0; i < parts; ++i) product_ell(y, x, A, lim[i], lim[i + 1]);
}

extern "C" {
void oracle_product_ell_mc_f(float *y, const float *x, ORACLE_ELL_ARGS(float), int parts) { ORACLE_MAKE_ELL(float); product_ell_mc(y, x, A, parts); }
void oracle_product_ell_mc_d(double *y, const double *x, ORACLE_ELL_ARGS(double), int parts) { ORACLE_MAKE_ELL(double); product_ell_mc(y, x, A, parts); }

void oracle_defect_ell_f(float *r, const float *rhs, const float *x, ORACLE_ELL_ARGS(float)) { ORACLE_MAKE_ELL(float); defect_ell(r, rhs, x, A); }
void oracle_defect_ell_d(double *r, const double *rhs, const double *x, ORACLE_ELL_ARGS(double)) { ORACLE_MAKE_ELL(double); defect_ell(r, rhs, x, A); }

// ---- BLAS-1 --------------------------------------------------------------------
float oracle_dot_f(const float *x, const float *y, u64 n) { return dot(x, y, n); }
double oracle_dot_d(const double *x, const double *y, u64 n) { return dot(x, y, n); }
float oracle_norm2_f(const float *x, u64 n) { return norm2(x, n); }
double oracle_norm2_d(const double *x, u64 n) { return norm2(x, n); }
void oracle_scaled_sum2_f(float *x, const float *y, float b, u64 n) { scaled_sum2(x, y, b, n); }
void oracle_scaled_sum2_d(double *x, const double *y, double b, u64 n) { scaled_sum2(x, y, b, n); }
void oracle_scaled_sum3_f(float *r, const float *x, const float *y, float b, u64 n) { scaled_sum3(r, x, y, b, n); }
void oracle_scaled_sum3_d(double *r, const double *x, const double *y, double b, u64 n) { scaled_sum3(r, x, y, b, n); }
void oracle_scaled_sum_v_f(float *a, const float *b, const float *c, u64 n) { scaled_sum_v(a, b, c, n); }
void oracle_scaled_sum_v_d(double *a, const double *b, const double *c, u64 n) { scaled_sum_v(a, b, c, n); }
void oracle_sum_f(float *x, const float *y, u64 n) { vsum(x, y, n); }
void oracle_sum_d(double *x, const double *y, u64 n) { vsum(x, y, n); }
void oracle_difference_f(float *x, const float *y, u64 n) { vdiff(x, y, n); }
void oracle_difference_d(double *x, const double *y, u64 n) { vdiff(x, y, n); }
void oracle_scale_f(float *x, float a, u64 n) { vscale(x, a, n); }
void oracle_scale_d(double *x, double a, u64 n) { vscale(x, a, n); }
void oracle_element_product_f(float *r, const float *x, const float *y, u64 n) { vprod(r, x, y, n); }
void oracle_element_product_d(double *r, const double *x, const double *y, u64 n) { vprod(r, x, y, n); }

// ---- Q1 operator -----------------------------------------------------------------
void oracle_q1_bands_f(u64 N, float *bands, int bc) { q1_bands(N, bands, bc); }
void oracle_q1_bands_d(u64 N, double *bands, int bc) { q1_bands(N, bands, bc); }
u64 oracle_banded_q1_to_csr_f(u64 N, const float *bands, u64 *rp, u64 *col, float *val) { return banded_q1_to_csr(N, bands, rp, col, val); }
u64 oracle_banded_q1_to_csr_d(u64 N, const double *bands, u64 *rp, u64 *col, double *val) { return banded_q1_to_csr(N, bands, rp, col, val); }
void oracle_banded_q1_f(u64 N, const float *bands, const float *rhs, const float *x, float *y) {
    banded_q1(N, bands, rhs, x, y);
}
void oracle_banded_q1_d(u64 N, const double *bands, const double *rhs, const double *x, double *y) {
    banded_q1(N, bands, rhs, x, y);
}
void oracle_q1_coeff_bands_f(u64 side, const double *kappa, float *bands) { q1_coeff_bands(side, kappa, bands); }
void oracle_q1_coeff_bands_d(u64 side, const double *kappa, double *bands) { q1_coeff_bands(side, kappa, bands); }
void oracle_coarsen_coeff(const double *fine, u64 nf, double *coarse, u64 nc, int averaging) {
    coarsen_coeff(fine, nf, coarse, nc, averaging);
}

// ---- transfers ---------------------------------------------------------------------
void oracle_restriction_f(float *c, u64 nc, const float *f, u64 nf, const u64 *mask) { restriction(c, nc, f, nf, mask); }
void oracle_restriction_d(double *c, u64 nc, const double *f, u64 nf, const u64 *mask) { restriction(c, nc, f, nf, mask); }
void oracle_prolongation_f(float *f, u64 nf, const float *c, u64 nc, const u64 *mask) { prolongation(f, nf, c, nc, mask); }
void oracle_prolongation_d(double *f, u64 nf, const double *c, u64 nc, const u64 *mask) { prolongation(f, nf, c, nc, mask); }
void oracle_apply_dirichlet_f(float *g, u64 N, const u64 *mask) { apply_dirichlet(g, N, mask); }
void oracle_apply_dirichlet_d(double *g, u64 N, const u64 *mask) { apply_dirichlet(g, N, mask); }
u64 oracle_prolmat_csr(u64 fine_dim, u64 *rp, u64 *col, double *val) { return prolmat_csr(fine_dim, rp, col, val); }

// ---- solvers -----------------------------------------------------------------------
void oracle_cg_solver_f(ORACLE_ELL_ARGS(float), const float *b, float *x, u64 max_iters, float eps, u64 *used, float *hist) { ORACLE_MAKE_ELL(float); cg_solver(A, b, x, max_iters, used, eps, hist); }
void oracle_cg_solver_d(ORACLE_ELL_ARGS(double), const double *b, double *x, u64 max_iters, double eps, u64 *used, double *hist) { ORACLE_MAKE_ELL(double); cg_solver(A, b, x, max_iters, used, eps, hist); }
void oracle_pcg_jacobi_f(ORACLE_ELL_ARGS(float), const float *b, float *x, const float *dinv, u64 max_iters, float eps, u64 *used, float *hist) { ORACLE_MAKE_ELL(float); pcg_jacobi(A, b, x, dinv, max_iters, used, eps, hist); }
void oracle_pcg_jacobi_d(ORACLE_ELL_ARGS(double), const double *b, double *x, const double *dinv, u64 max_iters, double eps, u64 *used, double *hist) { ORACLE_MAKE_ELL(double); pcg_jacobi(A, b, x, dinv, max_iters, used, eps, hist); }
u64 oracle_cg_konv_f(ORACLE_ELL_ARGS(float), const float *b, float *x, double konv) { ORACLE_MAKE_ELL(float); return cg_konv(A, b, x, konv); }
u64 oracle_cg_konv_d(ORACLE_ELL_ARGS(double), const double *b, double *x, double konv) { ORACLE_MAKE_ELL(double); return cg_konv(A, b, x, konv); }

// Multigrid<..., JAC, V, FIXED>::value on ELL (math/multigrid.hh:828-870):
// x[max] = b on the boundary rows/first column, then one kernel call with
// rhs[max] = b.
u64 oracle_mg_fixed_d(const oracle_mg_level *L, const oracle_mg_config *cfg, const double *b,
                      double *x, double *final_defect) {
    MGState<double> S;
    mg_alloc(L, *cfg, S);
    const u64 n = L[cfg->max_level].A.rows;
    mg_boundary_guess(x, b, n);
    vcopy(S.x[cfg->max_level].data(), x, n);
    vcopy(S.rhs[cfg->max_level].data(), b, n);
    const u64 it = mg_kernel(L, *cfg, S, final_defect);
    vcopy(x, S.x[cfg->max_level].data(), n);
    return it;
}
u64 oracle_mg_fixed_f(const oracle_mg_level *L, const oracle_mg_config *cfg, const float *b,
                      float *x, double *final_defect) {
    MGState<float> S;
    mg_alloc(L, *cfg, S);
    const u64 n = L[cfg->max_level].A.rows;
    mg_boundary_guess(x, b, n);
    vcopy(S.x[cfg->max_level].data(), x, n);
    vcopy(S.rhs[cfg->max_level].data(), b, n);
    const u64 it = mg_kernel(L, *cfg, S, final_defect);
    vcopy(x, S.x[cfg->max_level].data(), n);
    return it;
}

// Multigrid<..., JAC, V, MIXED>::value on ELL (math/multigrid.hh:2395-2536):
// FP32 levels in L, FP64 system A64.  Returns outer iterations; the sum of
// inner V-cycles goes to *inner_total; the FP64 relative defect after each
// outer step goes to hist[outer-1] when hist is given (capacity max_outer or
// 1000).
u64 oracle_mg_mixed(const oracle_mg_level *L, const oracle_mg_config *cfg,
                    const oracle_ell *A64e, const double *b, double *x,
                    u64 *inner_total, double *hist) {
    const u64 maxl = cfg->max_level;
    const u64 n = A64e->rows;
    Ell<double> A64 = as_ell<double>(*A64e);
    MGState<float> S;
    mg_alloc(L, *cfg, S);
    std::vector<double> result(n, 0.0), guess(x, x + n), product(n, 0.0),
        outer_defect(n), x_outer(n);
    mg_boundary_guess(guess.data(), b, n);
    for (u64 i = 0; i < n; ++i) S.x[maxl][i] = (float)guess[i];
    product_ell(product.data(), guess.data(), A64, 0, n);
    vcopy(outer_defect.data(), b, n);
    vdiff(outer_defect.data(), product.data(), n);
    const double initial_defect = std::sqrt(norm2(outer_defect.data(), n));
    double scale_factor = 1.0;
    u64 outer = 0, inner = 0;
    const u64 cap = cfg->max_outer ? cfg->max_outer : 1000;
    while (true) {
        for (u64 i = 0; i < n; ++i) S.rhs[maxl][i] = (float)outer_defect[i];
        if (cfg->zero_inner_start) vfill(S.x[maxl].data(), 0.0f, n);
        inner += mg_kernel(L, *cfg, S, nullptr);
        ++outer;
        for (u64 i = 0; i < n; ++i) x_outer[i] = (double)S.x[maxl][i];
        scaled_sum2(result.data(), x_outer.data(), scale_factor, n);
        defect_ell(outer_defect.data(), b, result.data(), A64);
        const double def_norm = std::sqrt(norm2(outer_defect.data(), n));
        if (hist && outer <= cap) hist[outer - 1] = def_norm / initial_defect;
        if (def_norm < cfg->tolerance * initial_defect) break;
        if (outer >= cap) break;
        scale_factor = def_norm;
        const double inv = 1.0 / scale_factor;
        vscale(outer_defect.data(), inv, n);
    }
    vcopy(x, result.data(), n);
    if (inner_total) *inner_total = inner;
    return outer;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Operator-list MG (math/mg.hh:547-1127): MGCycleCreation V/W/F STATIC with
// RISmoother (math/ri.hh:96-134, Jacobi preconditioner P = damping/diag),
// CGSolver<NONE> coarse solver, Restriction/Prolongation<PROLMAT>, driven by
// MGSolver (math/mg.hh:1093-1127).  Levels are numbered like the reference's
// base indices: A at levels 1..max, prol/res at levels 2..max.
// ---------------------------------------------------------------------------

namespace {

struct OpListMG {
    const oracle_mg_level *L;
    u64 min_level, max_iters_coarse, n_pre, n_post;
    double eps;
    int cycle;  // 0 = V, 1 = W, 2 = F
    std::vector<std::vector<double>> b, x, d, c, store;
    u64 used_coarse = 0;

    void smooth(u64 level, std::vector<double> &bb, std::vector<double> &xx, u64 iters) {
        Ell<double> A = as_ell<double>(L[level].A);
        const double *P = (const double *)L[level].dinv;
        const u64 n = A.rows;
        std::vector<double> t0(n), t1(n);
        for (u64 i = 0; i < iters; ++i) {
            defect_ell(t0.data(), bb.data(), xx.data(), A);
            vprod(t1.data(), P, t0.data(), n);
            vsum(xx.data(), t1.data(), n);
        }
    }
    void coarse(u64 level, std::vector<double> &bb, std::vector<double> &xx) {
        Ell<double> A = as_ell<double>(L[level].A);
        u64 used = 0;
        cg_solver(A, bb.data(), xx.data(), max_iters_coarse, &used, eps, (double *)nullptr);
        used_coarse = used;
    }
    // shape: 0 V, 1 W (second recursion unless the next level is coarsest),
    // F = F then V on the next level.
    void cyc(int shape, u64 level, std::vector<std::vector<double>> &X, std::vector<std::vector<double>> &B) {
        if (level == min_level) { coarse(level, B[level], X[level]); return; }
        Ell<double> A = as_ell<double>(L[level].A);
        smooth(level, B[level], X[level], n_pre);
        defect_ell(store[level].data(), B[level].data(), X[level].data(), A);
        {
            Ell<double> R = as_ell<double>(L[level].res);
            product_ell(d[level - 1].data(), store[level].data(), R, 0, R.rows);
        }
        vfill(c[level - 1].data(), 0.0, c[level - 1].size());
        if (shape == 0) {
            cyc(0, level - 1, c, d);
        } else if (shape == 1) {
            if (level - 1 != min_level) cyc(1, level - 1, c, d);
            cyc(1, level - 1, c, d);
        } else {
            cyc(2, level - 1, c, d);
            cyc(0, level - 1, c, d);
        }
        {
            Ell<double> P = as_ell<double>(L[level].prol);
            product_ell(store[level].data(), c[level - 1].data(), P, 0, P.rows);
        }
        vsum(X[level].data(), store[level].data(), X[level].size());
        smooth(level, B[level], X[level], n_post);
    }
};

}  // namespace

extern "C" {

// MGSolver<Tag, Norm<l2,true>>::value (math/mg.hh:1093-1127).  x (n_max) is
// the start vector on input and the solution on output.
u64 oracle_mg_oplist_d(const oracle_mg_level *L, u64 min_level, u64 max_level, int cycle,
                       u64 max_iters, u64 max_iters_coarse, u64 n_pre, u64 n_post, double eps,
                       const double *b, double *x, double *norms_out) {
    OpListMG mg;
    mg.L = L; mg.min_level = min_level; mg.max_iters_coarse = max_iters_coarse;
    mg.n_pre = n_pre; mg.n_post = n_post; mg.eps = eps; mg.cycle = cycle;
    const u64 nl = max_level + 1;
    mg.b.assign(nl, {}); mg.x.assign(nl, {}); mg.d.assign(nl, {}); mg.c.assign(nl, {}); mg.store.assign(nl, {});
    for (u64 l = 1; l <= max_level; ++l) {
        const u64 n = L[l].A.rows;
        mg.b[l].assign(n, 0.0); mg.x[l].assign(n, 0.0); mg.d[l].assign(n, 0.0);
        mg.c[l].assign(n, 0.0); mg.store[l].assign(n, 0.0);
    }
    const u64 n = L[max_level].A.rows;
    vcopy(mg.b[max_level].data(), b, n);
    vcopy(mg.x[max_level].data(), x, n);
    Ell<double> A = as_ell<double>(L[max_level].A);
    std::vector<double> r(n);
    defect_ell(r.data(), mg.b[max_level].data(), mg.x[max_level].data(), A);
    const double initial = std::sqrt(norm2(r.data(), n));
    double current = 1e16;
    u64 used = 0;
    for (u64 i = 0; i < max_iters; ++i) {
        mg.cyc(cycle, max_level, mg.x, mg.b);
        defect_ell(r.data(), mg.b[max_level].data(), mg.x[max_level].data(), A);
        current = std::sqrt(norm2(r.data(), n));
        used = i + 1;
        if (current < eps * initial) break;
    }
    vcopy(x, mg.x[max_level].data(), n);
    if (norms_out) { norms_out[0] = initial; norms_out[1] = current; }
    return used;
}

}  // extern "C"
