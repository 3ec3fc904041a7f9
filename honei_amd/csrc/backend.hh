// backend.hh -- internal C++ launch API of the MI355X backend.
//
// Every function enqueues work on `stream` and returns a hipError_t; none of
// them synchronises.  Device pointers are borrowed.  ELL matrices are in the
// reference's column-major layout (la/sparse_matrix_ell.hh:44-129): entry k
// of lane t of row r sits at slot r*T + t + k*stride, with 32-bit column
// indices and 32-bit row lengths (Arl, in units of T-wide chunks).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "hip_common.hh"

namespace honei {
namespace hipx {

enum ValueType { VT_F32 = 0, VT_F64 = 1, VT_BF16 = 2 };

struct EllView {
    int64_t rows, cols, stride, ncpr;
    int threads;
    const int32_t *Aj;
    const void *Ax;
    const uint32_t *Arl;
    ValueType vtype;
    // 32-bit words per stored index of Aj and Arl: 1 (int32 / uint32 arrays,
    // every kernel) or 2 (the reference's 64-bit arrays; ell_apply's product
    // and defect only)
    int index_words = 1;
};

// What the ELL row kernel does with the row sum s = sum_j A_rj x_j.
enum EllMode {
    ELL_PRODUCT = 0,   // out[r - row_start] = s
    ELL_DEFECT = 1,    // out[r] = rhs[r] - s            (+ ||out||^2 if norm)
    ELL_JACOBI = 2,    // out[r] = x[r] + dinv[r]*(rhs[r] - s)   (+ acc[r] += out[r])
    ELL_JACOBI0 = 3,   // as JACOBI with x = dinv .* rhs formed on the fly
};

struct EllArgs {
    EllMode mode = ELL_PRODUCT;
    void *out = nullptr;          // y / r / x_new (vector type)
    const void *x = nullptr;      // input vector (unused for JACOBI0)
    const void *rhs = nullptr;
    const void *dinv = nullptr;
    void *acc = nullptr;          // JACOBI*: acc += out when non-null
    double *norm2 = nullptr;      // DEFECT: ||out||^2 into this device double
    int64_t row_start = 0, row_end = 0;
    int block = 256;
    const int *gate = nullptr;    // non-null: the launch does nothing unless *gate != 0 (device flag)
};

// Vector element type follows the matrix: f32 and bf16 matrices act on f32
// vectors, f64 matrices on f64 vectors.
hipError_t ell_apply(const EllView &A, const EllArgs &a, const ReduceSlots &slots, hipStream_t s);

// Column window of every group of rows_per_block rows: ranges[2b], [2b+1] =
// smallest / largest stored column index of rows [b*rpb, (b+1)*rpb) (min >
// max when the rows hold no entry).  Depends on the pattern only.
hipError_t ell_block_ranges(const EllView &A, int64_t rows_per_block, int32_t *ranges, hipStream_t s);
// y = A x with each block's x window staged in LDS (`window` elements per
// work-group, at most kMaxStageBytes; wider windows are gathered from
// global memory); a work-group covers `chunks` x block/T rows and `ranges`
// come from ell_block_ranges with that rows_per_block.  Bit-identical to
// ell_apply's ELL_PRODUCT over all rows.
constexpr size_t kMaxStageBytes = 96 * 1024;
hipError_t ell_product_staged(const EllView &A, void *y, const void *x, const int32_t *ranges, int block, int chunks,
                              int64_t window, hipStream_t s);

// ---- BLAS-1 / reductions (double accumulation, deterministic order) -------
// Each reduction writes its result into a device double (out) and, when
// out_host_pinned is non-null, nowhere else; callers read it back when needed.
hipError_t dot(const void *x, const void *y, int64_t n, ValueType vt, double *out,
               const ReduceSlots &slots, hipStream_t s);
hipError_t norm2(const void *x, int64_t n, ValueType vt, double *out, const ReduceSlots &slots,
                 hipStream_t s);
// x += b*y ; b either a host value (b_dev null) or read from device memory;
// skip non-null: nothing happens while *skip != 0 (device flag).
hipError_t axpy(void *x, const void *y, double b, const double *b_dev, int64_t n, ValueType vt,
                hipStream_t s, const int *skip = nullptr);
// r = x + b*y
hipError_t scaled_sum3(void *r, const void *x, const void *y, double b, int64_t n, ValueType vt,
                       hipStream_t s);
// a += b .* c
hipError_t scaled_sum_v(void *a, const void *b, const void *c, int64_t n, ValueType vt, hipStream_t s);
hipError_t vsum(void *x, const void *y, int64_t n, ValueType vt, hipStream_t s);          // x += y
hipError_t vdiff2(void *x, const void *y, int64_t n, ValueType vt, hipStream_t s);        // x -= y
hipError_t vdiff3(void *r, const void *x, const void *y, int64_t n, ValueType vt, hipStream_t s);  // r = x - y
hipError_t vscale(void *x, double a, int64_t n, ValueType vt, hipStream_t s);             // x *= a
hipError_t vprod(void *r, const void *x, const void *y, int64_t n, ValueType vt, hipStream_t s);  // r = x.*y
hipError_t vinverse(void *x, int64_t n, ValueType vt, hipStream_t s);                      // x = 1/x
hipError_t vfill(void *x, double v, int64_t n, ValueType vt, hipStream_t s);
hipError_t vcopy(void *dst, const void *src, int64_t n, ValueType vt, hipStream_t s);
hipError_t convert_d2f(float *dst, const double *src, int64_t n, hipStream_t s);
hipError_t convert_f2d(double *dst, const float *src, int64_t n, hipStream_t s);
hipError_t fill_bytes_zero(void *p, size_t bytes, hipStream_t s);
// Bandwidth probe over n4 16-byte elements of src on `grid` 256-thread
// work-groups: kind 0 reads (one float per work-group into dst), kind 1
// copies into dst.
hipError_t stream_probe(int kind, void *dst, const void *src, int64_t n4, int grid, hipStream_t s);
// d = low 32 bits of the 64-bit indices src; *bad += entries >= bound
// (*bad is not cleared here).
hipError_t narrow_indices(uint32_t *d, const uint64_t *src, int64_t n, uint64_t bound, unsigned int *bad,
                          hipStream_t s);

// ---- sliced ELL (sell_kernels.hip) ----------------------------------------
// Rows sorted by length (longest first) within windows of sell_window()
// rows, in slices of sell_slice_rows(); see sell_kernels.hip.
int sell_window();
int sell_slice_rows();
hipError_t sell_sort(const uint32_t *Arl, int64_t rows, int32_t *perm, uint32_t *len, hipStream_t s);
hipError_t sell_widths(const uint32_t *len, int64_t rows, int64_t slices, int32_t *width, hipStream_t s);
hipError_t sell_fill(ValueType vt, const int32_t *Aj, const void *Ax, int64_t stride, int64_t rows,
                     const int32_t *perm, const uint32_t *len, const int32_t *width, const int64_t *base,
                     int32_t *sAj, void *sAx, hipStream_t s);
hipError_t sell_spmv(ValueType vt, const int32_t *sAj, const void *sAx, const int32_t *perm, const uint32_t *len,
                     const int32_t *width, const int64_t *base, int64_t rows, const void *x, void *y, hipStream_t s);
// Column range of each sort window's stored entries (ranges[2 w], [2 w + 1]).
hipError_t sell_ranges(const int32_t *sAj, const uint32_t *len, const int32_t *width, const int64_t *base,
                       int64_t rows, int32_t *ranges, hipStream_t s);
// sell_spmv with each window's x range staged in LDS (`window` entries of
// LDS per work-group, at most kMaxStageBytes); the same bits.
hipError_t sell_spmv_staged(ValueType vt, const int32_t *sAj, const void *sAx, const int32_t *perm,
                            const uint32_t *len, const int32_t *width, const int64_t *base, int64_t rows,
                            const void *x, void *y, const int32_t *ranges, int64_t window, hipStream_t s);

// Banded Q1 product y = A x (rhs null) or defect y = rhs - A x; bands[9] =
// LL, LD, LU, DL, DD, DU, UL, UD, UU device arrays of n = m*m values.
hipError_t q1_banded(ValueType vt, const void *const *bands, const void *x, const void *rhs, void *y, int64_t n,
                     int64_t m, int block, hipStream_t s);
// x += alpha*y and return ||x||^2 (fused; alpha from device memory when
// non-null); skip as for axpy (out then keeps its value).
hipError_t axpy_norm2(void *x, const void *y, double alpha, const double *alpha_dev, int64_t n,
                      ValueType vt, double *out, const ReduceSlots &slots, hipStream_t s,
                      const int *skip = nullptr);

// ---- geometric transfers on (2^L+1)^2 lexicographic grids --------------
// mask: 8 entries (corners then edges) on the host; value 2 = Dirichlet.
hipError_t restriction(void *coarse, int64_t n_coarse, const void *fine, int64_t n_fine,
                       const uint64_t *mask, ValueType vt, hipStream_t s);
hipError_t prolongation(void *fine, int64_t n_fine, const void *coarse, int64_t n_coarse,
                        const uint64_t *mask, ValueType vt, hipStream_t s);
// fine = P coarse (+ Dirichlet mask); x += alpha*fine fused when x non-null.
hipError_t prolongation_axpy(void *fine, int64_t n_fine, const void *coarse, int64_t n_coarse,
                             const uint64_t *mask, void *x, double alpha, ValueType vt, hipStream_t s);

}  // namespace hipx
}  // namespace honei
