/*
 * honei_hip.h -- C ABI of the MI355X (gfx950) backend for HONEI's ELL
 * sparse-solve path.
 *
 * Plain pointers and sizes only.  Pointers named *_dev or documented as
 * device pointers are device memory obtained from honei_hip_malloc (or any
 * hipMalloc); `stream` is a hipStream_t passed as void* (NULL = default
 * stream).  Every function returns an int status (HONEI_HIP_SUCCESS = 0);
 * no C++ exception crosses this boundary.  honei_hip_last_error() describes
 * the last failure on the calling thread.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the HONEI source tree).  Differences in argument
 * convention are stated next to the declaration.
 */
#ifndef HONEI_HIP_H
#define HONEI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status -------------------------------------------------------------- */
#define HONEI_HIP_SUCCESS 0
#define HONEI_HIP_ERROR_INVALID_ARGUMENT 1 /* bad pointer / size / layout       */
#define HONEI_HIP_ERROR_SIZE_MISMATCH 2    /* VectorSizeDoesNotMatch analogue    */
#define HONEI_HIP_ERROR_OUT_OF_MEMORY 3
#define HONEI_HIP_ERROR_RUNTIME 4          /* HIP runtime / launch failure       */
#define HONEI_HIP_ERROR_NOT_CONVERGED 5    /* iteration limit reached            */
#define HONEI_HIP_ERROR_BUDGET 6           /* request exceeds the memory budget  */
#define HONEI_HIP_ERROR_NO_DEVICE 7

/* value/vector type codes used by the generic entry points */
#define HONEI_HIP_F32 0
#define HONEI_HIP_F64 1
#define HONEI_HIP_BF16 2 /* matrix values only; vectors are f32 */

const char *honei_hip_last_error(void);
const char *honei_hip_version(void);

/* ---- device and memory -------------------------------------------------
 * Replace backends/cuda/multi_gpu.cu:23-71 (device count/select),
 * backends/cuda/operations.cu:22-26 (cuda_thread_synchronize) and
 * backends/cuda/transfer.cu:49-160 (cuda_malloc, cuda_upload, cuda_download,
 * cuda_copy, cuda_free, cuda_fill_zero, cuda_convert_*).  The fill is a
 * kernel, not hipMemset, so it may be captured in a hipGraph.            */
int honei_hip_device_count(int *count);
int honei_hip_init(int device);
int honei_hip_shutdown(void);
int honei_hip_malloc(void **ptr, size_t bytes);
/* Pinned host memory (asynchronous copies); replaces cuda_malloc_host /
 * cuda_free_host (backends/cuda/transfer.cu:49-60).                     */
int honei_hip_host_alloc(void **ptr, size_t bytes);
int honei_hip_host_free(void *ptr);
int honei_hip_free(void *ptr);
int honei_hip_upload(void *dst_dev, const void *src_host, size_t bytes, void *stream);
int honei_hip_download(void *dst_host, const void *src_dev, size_t bytes, void *stream);
int honei_hip_copy(void *dst_dev, const void *src_dev, size_t bytes, void *stream);
int honei_hip_fill_zero(void *dst_dev, size_t bytes, void *stream);
int honei_hip_fill(int vtype, void *x_dev, double value, size_t n, void *stream); /* fill<Tag>(x, value) */
int honei_hip_synchronize(void *stream);
int honei_hip_stream_create(void **stream);
int honei_hip_stream_destroy(void *stream);
/* HIP events on a given stream, for timing launches where they run.       */
int honei_hip_event_create(void **event);
int honei_hip_event_destroy(void *event);
int honei_hip_event_record(void *event, void *stream);
int honei_hip_event_elapsed_ms(float *ms, void *start, void *stop); /* waits for stop */
int honei_hip_convert_double_float(float *dst_dev, const double *src_dev, size_t n, void *stream);
int honei_hip_convert_float_double(double *dst_dev, const float *src_dev, size_t n, void *stream);

/* ELL index upload.  The reference keeps Aj/Arl as 64-bit unsigned long on
 * the device (la/sparse_matrix_ell.hh:110-116); this backend stores 32-bit
 * indices.  Narrows on the host, checks every value is < bound (columns for
 * Aj; pass UINT64_MAX to skip) and uploads.  Synchronous.                 */
int honei_hip_upload_indices(void *dst_dev_u32, const uint64_t *src_host, size_t n, uint64_t bound);
/* Matrix values stored as bf16 (round to nearest even) for the optional
 * reduced-storage FP32 path.  Synchronous.                                 */
int honei_hip_upload_values_bf16(void *dst_dev_bf16, const float *src_host, size_t n);

/* ---- ELL operators -----------------------------------------------------
 * Aj: int32 [ncpr*stride], Arl: uint32 [rows], Ax: values [ncpr*stride];
 * threads (T) in {1,2,4,8,16,32,64} must equal the matrix layout's T.
 * blocksize 0 = default (256), else a multiple of 64 in [T, 1024].
 *
 * Replaces cuda_product_smell_dv_{float,double}
 * (backends/cuda/operations.hh:87-93): y[r - row_start] = (A x)[r] for
 * r in [row_start, row_end), the reference's sub-range convention
 * (backends/cuda/product.cu:194).  _bf16: Ax bf16, x/y float.            */
int honei_hip_product_smell_dv_float(void *x, void *y, const void *Aj, const void *Ax, const void *Arl,
                                     size_t row_start, size_t row_end, size_t num_cols_per_row, size_t stride,
                                     size_t blocksize, size_t threads, void *stream);
int honei_hip_product_smell_dv_double(void *x, void *y, const void *Aj, const void *Ax, const void *Arl,
                                      size_t row_start, size_t row_end, size_t num_cols_per_row, size_t stride,
                                      size_t blocksize, size_t threads, void *stream);
int honei_hip_product_smell_dv_bf16(void *x, void *y, const void *Aj, const void *Ax, const void *Arl,
                                    size_t row_start, size_t row_end, size_t num_cols_per_row, size_t stride,
                                    size_t blocksize, size_t threads, void *stream);

/* ---- ELL product with x staged in LDS (new) --------------------------------
 * Product<tags::GPU::HIP> on SparseMatrixELL for matrices whose rows keep
 * their columns near the diagonal (banded / FEM orderings, la/product.hh:
 * 148-206 computes the same sums): each work-group copies the x window of
 * its rows into LDS once and gathers from there instead of from L2.
 * honei_hip_ell_block_ranges: per group of rows_per_block rows the smallest
 * and largest stored column other than 0 (padding slots hold column 0; x[0]
 * is read separately) as int32 pairs, device, [ceil(rows/rpb)][2], and,
 * when max_width is non-null, the widest window (host; synchronises).  It
 * depends on the pattern only: compute once per matrix.
 * honei_hip_product_smell_dv_staged: y = A x over all rows; a work-group
 * covers chunks (1..64) x blocksize/threads rows and the ranges are made
 * with that rows_per_block (blocksize 0 = 256; nranges = pairs in `ranges`,
 * at least the number of work-groups, else an error); window = LDS
 * elements per work-group (1 .. 96 KiB worth; use max_width when it fits).
 * A block whose window is wider, or a column outside its range, is gathered
 * from global memory.  Bit-identical to honei_hip_product_smell_dv_*.     */
int honei_hip_ell_block_ranges(const void *Aj, const void *Arl, size_t rows, size_t num_cols_per_row, size_t stride,
                               size_t threads, size_t rows_per_block, int32_t *ranges, int64_t *max_width,
                               void *stream);
int honei_hip_product_smell_dv_staged(int vtype, void *x, void *y, const void *Aj, const void *Ax, const void *Arl,
                                      size_t rows, size_t num_cols_per_row, size_t stride, size_t threads,
                                      size_t blocksize, size_t chunks, const int32_t *ranges, size_t nranges,
                                      int64_t window, void *stream);

/* ---- banded Q1 operators ---------------------------------------------------
 * BandedMatrixQx<Q1Type, DT>: nine device bands of `size` = m*m values
 * (LL, LD, LU, DL, DD, DU, UL, UD, UU: coefficients towards columns i-m-1,
 * i-m, i-m+1, i-1, i, i+1, i+m-1, i+m, i+m+1).  blocksize 0 = 256, else a
 * multiple of 64 up to 256.  Each row is summed in the order of
 * Product<tags::CPU>::value(BandedMatrixQx, DenseVector) (la/product.hh:
 * 371-400), so results are bit-identical to the CPU path.
 * Replaces cuda_product_bmdv_q1_{float,double}
 * (backends/cuda/operations.hh:77-85): y = A x, and cuda_defect_q1_{float,
 * double} (backends/cuda/operations.hh:31-39): y = rhs - A x.            */
int honei_hip_product_bmdv_q1_float(void *ll, void *ld, void *lu, void *dl, void *dd, void *du, void *ul, void *ud,
                                    void *uu, void *x, void *y, size_t size, size_t blocksize, size_t m, void *stream);
int honei_hip_product_bmdv_q1_double(void *ll, void *ld, void *lu, void *dl, void *dd, void *du, void *ul, void *ud,
                                     void *uu, void *x, void *y, size_t size, size_t blocksize, size_t m,
                                     void *stream);
int honei_hip_defect_q1_float(void *rhs, void *ll, void *ld, void *lu, void *dl, void *dd, void *du, void *ul,
                              void *ud, void *uu, void *x, void *y, size_t size, size_t blocksize, size_t m,
                              void *stream);
int honei_hip_defect_q1_double(void *rhs, void *ll, void *ld, void *lu, void *dl, void *dd, void *du, void *ul,
                               void *ud, void *uu, void *x, void *y, size_t size, size_t blocksize, size_t m,
                               void *stream);

/* Replaces cuda_defect_smell_dv_{float,double}
 * (backends/cuda/operations.hh:41-47): result[r] = rhs[r] - (A b)[r]; same
 * argument order (rhs, result, Aj, Ax, Arl, b).                            */
int honei_hip_defect_smell_dv_float(void *rhs, void *result, const void *Aj, const void *Ax, const void *Arl,
                                    void *b, size_t row_start, size_t row_end, size_t num_cols_per_row,
                                    size_t stride, size_t blocksize, size_t threads, void *stream);
int honei_hip_defect_smell_dv_double(void *rhs, void *result, const void *Aj, const void *Ax, const void *Arl,
                                     void *b, size_t row_start, size_t row_end, size_t num_cols_per_row,
                                     size_t stride, size_t blocksize, size_t threads, void *stream);

/* ---- the reference's 64-bit index arrays ----------------------------------
 * SparseMatrixELL keeps Aj and Arl as unsigned long (la/sparse_matrix_ell.hh:
 * 110-116) and its CUDA glue passes their device copies straight to the
 * launcher (la/product-cuda.cc:178-194 for Product, math/defect-cuda.cc:
 * 524-660 for Defect).  These entries take those arrays as they are: Aj
 * uint64 [ncpr*stride], Arl uint64 [rows], column indices below 2^31 (the
 * low word of each index is read).  Otherwise identical to the int32 forms
 * above (same arguments, sub-range convention and bits); they stream 8 bytes
 * per index slot instead of 4.
 * Replace cuda_product_smell_dv_{float,double} and
 * cuda_defect_smell_dv_{float,double} (backends/cuda/operations.hh:41-47,
 * 87-93) for a caller that keeps the reference's index type.             */
int honei_hip_product_smell_dv_float_idx64(void *x, void *y, const void *Aj, const void *Ax, const void *Arl,
                                           size_t row_start, size_t row_end, size_t num_cols_per_row, size_t stride,
                                           size_t blocksize, size_t threads, void *stream);
int honei_hip_product_smell_dv_double_idx64(void *x, void *y, const void *Aj, const void *Ax, const void *Arl,
                                            size_t row_start, size_t row_end, size_t num_cols_per_row,
                                            size_t stride, size_t blocksize, size_t threads, void *stream);
int honei_hip_defect_smell_dv_float_idx64(void *rhs, void *result, const void *Aj, const void *Ax, const void *Arl,
                                          void *b, size_t row_start, size_t row_end, size_t num_cols_per_row,
                                          size_t stride, size_t blocksize, size_t threads, void *stream);
int honei_hip_defect_smell_dv_double_idx64(void *rhs, void *result, const void *Aj, const void *Ax,
                                           const void *Arl, void *b, size_t row_start, size_t row_end,
                                           size_t num_cols_per_row, size_t stride, size_t blocksize, size_t threads,
                                           void *stream);
/* Device-side narrowing of a 64-bit index array (the device copy of
 * SparseMatrixELL::Aj() or Arl()) into this backend's 32-bit form:
 * dst_dev_u32[i] = src_dev_u64[i], every value checked < bound (columns for
 * Aj, num_cols_per_row + 1 for Arl; values at or above 2^32 always fail).
 * For glue that narrows once per matrix and then calls the int32 entries
 * (4 instead of 8 bytes per index slot).  Synchronous (reads back the
 * count of out-of-range values); no reference counterpart.               */
int honei_hip_narrow_indices(void *dst_dev_u32, const void *src_dev_u64, size_t n, uint64_t bound, void *stream);

/* ---- sliced copy of an ELL matrix (no reference counterpart) ---------------
 * For matrices whose row lengths vary (ELL pads every row to the widest):
 * honei_hip_sell_create builds, on the device, a copy of a threads = 1 ELL
 * matrix (FP32 or FP64 values, int32 Aj, uint32 Arl) whose rows are sorted by
 * length, longest first, within windows of 2048 rows and grouped into slices
 * of 64 rows padded to their own longest row; each row keeps its entries in
 * stored order.  honei_hip_product_sell computes y = A x over all rows with
 * each row's terms added in that order: bit-identical to
 * honei_hip_product_smell_dv_* and Product<tags::CPU> (la/product.hh:
 * 148-206).  The handle owns its device memory (create allocates and
 * synchronises; product neither allocates nor synchronises).  create also
 * records each 2048-row window's column range; where it fits 96 KiB of LDS
 * the product copies that range of x into LDS once per window and gathers
 * from there (the same bits); honei_hip_sell_stage_x(sell, 0) turns this
 * off (gathers from global x), 1 back on (default).                       */
int honei_hip_sell_create(void **sell, int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                          size_t cols, size_t num_cols_per_row, size_t stride, size_t threads, void *stream);
int honei_hip_sell_info(void *sell, size_t *slices, size_t *padded_slots, size_t *device_bytes);
int honei_hip_product_sell(void *sell, void *x, void *y, void *stream);
int honei_hip_sell_stage_x(void *sell, int on);
int honei_hip_sell_destroy(void *sell);

/* Bandwidth probe (measurement only, no reference counterpart): `bytes`
 * (multiple of 16, 16-byte aligned pointers) streamed with 16-byte
 * non-temporal loads, eight in flight per thread, on `grid` (1..65536)
 * 256-thread work-groups.  kind 0: read src, one float sum per work-group
 * into dst (dst holds `grid` floats); kind 1: copy src to dst; both
 * grid-stride.  kinds 2 / 3: the same, each work-group taking one
 * contiguous chunk, chunks contiguous per XCD (as the SpMV's row blocks). */
int honei_hip_stream_probe(int kind, void *dst, const void *src, size_t bytes, size_t grid, void *stream);

/* Fused defect + norm (no reference counterpart; the reference runs
 * Defect then Norm, math/multigrid.hh:2495-2498): result = rhs - A x over
 * all rows and *norm2_dev = ||result||^2 (double accumulation), in one pass. */
int honei_hip_defect_norm_smell_dv(int vtype, void *rhs, void *result, const void *Aj, const void *Ax,
                                   const void *Arl, void *x, size_t rows, size_t num_cols_per_row, size_t stride,
                                   size_t threads, double *norm2_dev, void *stream);

/* Fused Jacobi sweep (math/jacobi.hh:268-279 = Defect + ScaledSum in one
 * pass): x_new = x + dinv .* (rhs - A x); when acc is non-NULL also
 * acc += x_new.  x_new must not alias x.  With x == NULL the input iterate
 * is dinv .* rhs formed on the fly (first sweep from zero,
 * math/jacobi.hh:481-493, fused with the second).                          */
int honei_hip_jacobi_smell_dv(int vtype, void *x_new, const void *x, const void *rhs, const void *dinv,
                              const void *Aj, const void *Ax, const void *Arl, size_t rows,
                              size_t num_cols_per_row, size_t stride, size_t threads, void *acc, void *stream);

/* ---- BLAS-1 --------------------------------------------------------------
 * Reductions: cuda_dot_product_two_* and cuda_norm_l2_one_*
 * (backends/cuda/operations.hh:60-75) return the value on the host after a
 * device-side final sum (no partial-sum download); the _dev forms leave it
 * in device memory (double) without synchronising.  norm_l2_one is the
 * squared norm, as Norm<vnt_l_two, false>.                                 */
int honei_hip_dot_product_two_float(float *result, const void *x, const void *y, size_t n, void *stream);
int honei_hip_dot_product_two_double(double *result, const void *x, const void *y, size_t n, void *stream);
int honei_hip_norm_l2_one_float(float *result, const void *x, size_t n, void *stream);
int honei_hip_norm_l2_one_double(double *result, const void *x, size_t n, void *stream);
int honei_hip_dot_product_dev(int vtype, double *result_dev, const void *x, const void *y, size_t n, void *stream);
int honei_hip_norm_l2_dev(int vtype, double *result_dev, const void *x, size_t n, void *stream);
/* Fused axpy + norm (new): x += alpha*y; *result_dev = ||x||^2.            */
int honei_hip_axpy_norm_dev(int vtype, void *x, const void *y, double alpha, double *result_dev, size_t n,
                            void *stream);

/* Replaces cuda_scaled_sum_three_{float,double}_s (r = x + s*y),
 * cuda_scaled_sum_three_{float,double} (x += y .* z),
 * cuda_sum_two_*, cuda_difference_two_* / _three_*, cuda_scale_one_*,
 * cuda_element_product_three_*, cuda_element_inverse_one_*
 * (backends/cuda/operations.hh:55-69,119-128); plus the two-operand axpy
 * x += s*y used by ScaledSum (la/scaled_sum.hh:73-88).                     */
int honei_hip_scaled_sum_two(int vtype, void *x, const void *y, double s, size_t n, void *stream);
int honei_hip_scaled_sum_three_s(int vtype, void *r, const void *x, const void *y, double s, size_t n,
                                 void *stream);
int honei_hip_scaled_sum_three(int vtype, void *x, const void *y, const void *z, size_t n, void *stream);
int honei_hip_sum_two(int vtype, void *a, const void *b, size_t n, void *stream);
int honei_hip_difference_two(int vtype, void *a, const void *b, size_t n, void *stream);
int honei_hip_difference_three(int vtype, void *r, const void *a, const void *b, size_t n, void *stream);
int honei_hip_scale_one(int vtype, void *x, double a, size_t n, void *stream);
int honei_hip_element_product_three(int vtype, void *r, const void *a, const void *b, size_t n, void *stream);
int honei_hip_element_inverse_one(int vtype, void *x, size_t n, void *stream);

/* ---- geometric transfers -------------------------------------------------
 * Replace cuda_restriction_* / cuda_prolongation_*
 * (backends/cuda/operations.hh:101-111) on lexicographic (2^L+1)^2 grids.
 * mask: 8 host values (corners BL, BR, TR, TL, then edges B, R, T, L;
 * 2 = Dirichlet), as MGInfo::macro_border_mask (math/multigrid.hh:78).
 * Results are bit-identical to the CPU path.                              */
int honei_hip_restriction(int vtype, void *coarse, size_t size_coarse, const void *fine, size_t size_fine,
                          const uint64_t *mask, void *stream);
int honei_hip_prolongation(int vtype, void *fine, size_t size_fine, const void *coarse, size_t size_coarse,
                           const uint64_t *mask, void *stream);

/* ---- device Q1 assembly ----------------------------------------------------
 * Replaces the host chain FillMatrix<POISSON, DIRICHLET_0>
 * (math/fill_matrix.hh:36-78) -> SparseMatrix(BandedMatrixQx)
 * (la/sparse_matrix.hh:221-300) -> SparseMatrixELL with `threads` = T
 * (la/sparse_matrix_ell-impl.hh:75-122), built directly in HBM: Aj int32 and
 * Ax [ceil(9/T)*stride], Arl uint32 [side^2], stride >= side^2*T and a
 * multiple of T (reference: 32*ceil(side^2*T/32)); T a power of two <= 32.
 * coeff_dev NULL: the reference's constant stencil, bit-identical arrays.
 * coeff_dev non-NULL: per-cell coefficient (double, (side-1)^2, row-major
 * cells) -- the variable-coefficient Q1 operator, which the reference lacks.
 * dinv_dev (optional, vector type) = omega / diag, as the reference tests
 * build it (math/poisson_mg_fixed_ell_TEST.cc:707-714).                    */
int honei_hip_q1_assemble(int vtype, size_t side, void *Aj, void *Ax, void *Arl, size_t stride,
                          const double *coeff_dev, void *dinv_dev, double omega, size_t threads,
                          void *stream);

/* ---- synthetic workloads -----------------------------------------------------
 * Random ELL of the SpMV benchmark configuration generated in HBM (no
 * reference counterpart; the reference reads matrices from files): T = 1,
 * row lengths U{1..max_len}, distinct sorted columns uniform in
 * [r-band, r+band] clipped to [0, n), values N(0,1), in the T-interleaved
 * layout of `threads` (stride >= n*threads; Aj/Ax hold
 * ceil(max_len/threads)*stride slots, zero padded).  Deterministic in
 * (n, max_len, band, seed); the row contents do not depend on threads.     */
int honei_hip_random_ell(int vtype, size_t n, size_t max_len, size_t band, uint64_t seed, void *Aj, void *Ax,
                         void *Arl, size_t stride, size_t threads, void *stream);

/* ---- CG --------------------------------------------------------------------
 * CGSolver<Tag, methods::NONE>::value (math/cg.hh:63-133) on a device ELL
 * matrix: x (device, vector type) is the start vector and the result.
 * Scalars stay on the device; the host reads the convergence flag once per
 * chunk of iterations.  Returns HONEI_HIP_ERROR_NOT_CONVERGED when max_iters
 * is reached first (x and *used_iters are still written).                  */
int honei_hip_cg_solve(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                       size_t num_cols_per_row, size_t stride, size_t threads, const void *b_dev, void *x_dev,
                       size_t max_iters, double eps_relative, size_t *used_iters, double *final_defect,
                       void *stream);

/* honei_hip_cg_solve without allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory at call time (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels, results and status codes as honei_hip_cg_solve;
 * the solve's scratch (reduction partials, scalars, vectors p, r, v) is a
 * fixed library-held device array of 4 MiB + 4 KiB + 32 MiB per device,
 * reserved when the library is loaded and kept until the process ends.
 * rows is limited to what three vectors fit in 32 MiB (1,398,080 rows for
 * F64 and 2,796,160 for F32 / BF16, each vector rounded up to 256 bytes);
 * larger systems return HONEI_HIP_ERROR_INVALID_ARGUMENT without running.
 * Every call on one device uses the same scratch: calls that overlap on one
 * device (other threads or streams) race and must be serialised.          */
int honei_hip_cg_solve_noalloc(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                               size_t num_cols_per_row, size_t stride, size_t threads, const void *b_dev,
                               void *x_dev, size_t max_iters, double eps_relative, size_t *used_iters,
                               double *final_defect, void *stream);

/* ConjugateGradients<Tag, methods::JAC>::value(A, b, x, dd_inverted,
 * max_iters, used_iters, eps_relative) on ELL
 * (math/conjugate_gradients.hh:1110-1154, kernel :884-913): Jacobi-
 * preconditioned CG, z = dd_inverted .* r, stopping once
 * ||r|| < ||r_0|| * eps_relative.  dd_inverted is a device vector of the
 * vector type (double for F64 matrices, float otherwise).  *used_iters is
 * the iteration count (0 for max_iters = 0); *final_defect the last ||r||.
 * Returns HONEI_HIP_ERROR_NOT_CONVERGED when max_iters is reached first (x
 * and *used_iters are still written; the reference returns normally then). */
int honei_hip_pcg_jacobi(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                         size_t num_cols_per_row, size_t stride, size_t threads, const void *b_dev, void *x_dev,
                         const void *dd_inverted_dev, size_t max_iters, double eps_relative, size_t *used_iters,
                         double *final_defect, void *stream);

/* ConjugateGradients<Tag, methods::NONE>::value(A, b, x, konv_rad)
 * (math/conjugate_gradients.hh:154-201, 587-623), the multigrid coarse
 * solver: x = 0, iterate until ||x||^2 grows by no more than konv_rad.
 * Runs inside one work-group (for the small coarse grids it serves);
 * max_iters bounds the loop (the reference has no bound).                   */
int honei_hip_cg_konv(int vtype, const void *Aj, const void *Ax, const void *Arl, size_t rows,
                      size_t num_cols_per_row, size_t stride, size_t threads, const void *b_dev, void *x_dev,
                      double konv_rad, size_t max_iters, size_t *iters, void *stream);

/* ---- multigrid ---------------------------------------------------------------
 * A device-resident hierarchy plays the role of MGInfo
 * (math/multigrid.hh:66-149); solve_fixed / solve_mixed are
 * Multigrid<..., JAC, CYCLE::V, FIXED|MIXED>::value on ELL
 * (math/multigrid.hh:828-870, 2395-2536) with the reference's control flow.
 * Levels are numbered like the reference (level l = (2^l+1)^2 nodes for the
 * geometric transfers).  Host arrays are uploaded (indices narrowed to 32
 * bits); build_q1 assembles every level and the FP64 outer system on the
 * device instead.                                                          */
typedef struct honei_hip_mg_config {
    size_t n_max_iter;            /* V-cycles per inner solve             */
    size_t n_pre_smooth;          /* Jacobi sweeps (omega inside dinv)    */
    size_t n_post_smooth;         /* used only if post_smooth_uses_pre==0 */
    int initial_zero;
    int convergence_check;
    int prolmat;                  /* 1: transfers are ELL products P / R  */
    int zero_inner_start;         /* 1: inner start vector 0 every outer step */
    double tolerance;             /* relative, inner and outer            */
    double adapt_correction_factor; /* 0: alpha = (d,c)/(Ac,c)            */
    uint64_t mask[8];             /* macro_border_mask                    */
    size_t max_outer;             /* 0: no limit besides a 10000 cap      */
    int post_smooth_uses_pre;     /* 1: reference behaviour               */
} honei_hip_mg_config;

int honei_hip_mg_create(void **mg, int inner_vtype, size_t min_level, size_t max_level);
int honei_hip_mg_destroy(void *mg);
int honei_hip_mg_configure(void *mg, const honei_hip_mg_config *cfg);
int honei_hip_mg_set_level(void *mg, size_t level, size_t rows, size_t stride, size_t num_cols_per_row,
                           size_t threads, const uint64_t *Aj, const void *Ax, const uint64_t *Arl,
                           const void *dinv);
/* kind 0: prolongation matrix into `level`; kind 1: restriction from it.   */
int honei_hip_mg_set_transfer(void *mg, size_t level, int kind, size_t rows, size_t cols, size_t stride,
                              size_t num_cols_per_row, size_t threads, const uint64_t *Aj, const void *Ax,
                              const uint64_t *Arl);
int honei_hip_mg_set_outer(void *mg, size_t rows, size_t stride, size_t num_cols_per_row, size_t threads,
                           const uint64_t *Aj, const double *Ax, const uint64_t *Arl);
int honei_hip_mg_build_q1(void *mg, double omega, void *stream);
/* Variable-coefficient Q1 hierarchy (no reference counterpart): coeff_dev
 * holds (2^max_level)^2 per-cell values (double, row-major); coarser levels
 * average the four fine cells (averaging 0 arithmetic, 1 geometric,
 * 2 harmonic) and are rediscretised.                                        */
int honei_hip_mg_build_q1_coeff(void *mg, double omega, const double *coeff_dev, int averaging, void *stream);
int honei_hip_mg_solve_fixed(void *mg, const void *b_dev, void *x_dev, size_t *iters, double *final_defect,
                             void *stream);
int honei_hip_mg_solve_mixed(void *mg, const double *b_dev, double *x_dev, size_t *outer_iters,
                             size_t *inner_iters, double *rel_defect, void *stream);
/* FP64 conjugate gradients on the outer system, preconditioned by one lean
 * V-cycle of the hierarchy per iteration (z = |r| V(r/|r|); any inner type,
 * FP64 inner = the plain FP64 solve).  No reference counterpart: the
 * reference's PCG variants (math/conjugate_gradients.hh:884-913) use a
 * Jacobi preconditioner.  x_dev: start vector (boundary values as the MIXED
 * driver) and result; cfg.tolerance on |r|/|r0|, cfg.max_outer caps the
 * iterations (0: 10000); converged only when a recomputed b - A x confirms.
 * HONEI_HIP_ERROR_NOT_CONVERGED if the cap or a breakdown stops it.        */
int honei_hip_mg_solve_pcg(void *mg, const double *b_dev, double *x_dev, size_t *iters, double *rel_defect,
                           void *stream);

/* ---- seeded fields (synthetic Poisson workloads) ------------------------------
 * No reference counterpart.  Element i is drawn from position i of the
 * SplitMix64 sequence of `seed` (U = (u >> 11) * 2^-53):
 * uniform: dst[i] = lo + (hi - lo) * U_i;
 * lognormal: dst[i] = exp(sigma * z_i), z_i = sqrt(-2 ln(1 - U_2i)) * cos(2 pi U_2i+1). */
int honei_hip_fill_uniform(double *dst_dev, size_t n, double lo, double hi, uint64_t seed, void *stream);
int honei_hip_fill_lognormal(double *dst_dev, size_t n, double sigma, uint64_t seed, void *stream);

/* ---- batched multigrid ----------------------------------------------------------
 * No reference counterpart: the reference solves one system per call
 * (Multigrid<...MIXED>::value, math/multigrid.hh:2395-2536).  A batch is a
 * device-resident pool of slots in classes by finest level (side 2^L + 1,
 * L in 2..13, at most 16 classes, 65535 slots); every slot holds a Q1
 * hierarchy with a per-cell coefficient down to the 3x3 grid, solved by
 * the MIXED defect correction (FP64 outer defect, update and norm; FP32 or
 * BF16-stored V-cycle with n_pre / n_post damped Jacobi sweeps and the
 * single-work-group CG on the coarsest grid) in an FP64 outer iteration
 * (defect correction or V-cycle-preconditioned CG).  One graph launch runs
 * iters_per_launch outer iterations of every active slot; slots retire on
 * the device when ||b - A x|| < tolerance * ||b - A x0||, where x0 is zero
 * with the first row, last row and first column of the grid set to b.
 * autotune: threads-per-row T in {1,2,4} and block size per level picked by
 * timing (cached for the process); 0 uses T = 1, block 256, which with
 * outer_method 0 makes each slot bit-identical to honei_hip_mg_solve_mixed
 * with n_max_iter 1, initial_zero, zero_inner_start and post sweeps = pre
 * sweeps.  The batch
 * owns its streams; calls on one batch must not overlap.                    */
typedef struct {
    int32_t inner_vtype;       /* HONEI_HIP_F32 or HONEI_HIP_BF16 matrix values with FP32
                                  vectors (mixed precision), or HONEI_HIP_F64 (values and
                                  vectors in FP64: the same solver without the FP32 inner) */
    int32_t n_pre, n_post;     /* Jacobi sweeps, n_pre >= 2 */
    double omega;              /* Jacobi damping */
    double tolerance;          /* relative FP64 defect */
    int32_t max_outer;         /* outer iterations before a slot gives up */
    int32_t averaging;         /* coefficient coarsening: 0 arithmetic, 1 geometric, 2 harmonic */
    int32_t iters_per_launch;  /* outer iterations per graph launch (1..64, default 2) */
    int32_t autotune;
    int32_t outer_method;      /* 0 defect correction x += V(b - A x) (the reference's MIXED
                                  driver); 1 FP64 conjugate gradients preconditioned by the V-cycle,
                                  converged residual confirmed by a recomputed b - A x (default) */
    int32_t fused_levels;      /* levels 1..fused_levels (0..6, default 4) of every V-cycle run
                                  as one launch, one work-group per slot (same arithmetic) */
    int32_t use_graphs;        /* 1 (default): a step is one hipGraph launch; 0: the same kernels
                                  enqueued directly at every step (same results; for profilers
                                  that cannot follow graph launches) */
    int32_t pair_sweeps;       /* 1 (default): on levels of 33^2 nodes and up, Jacobi0 + defect and
                                  the two post-smoothing sweeps each run as one launch that reads
                                  the matrix once (n_pre == 2, n_post == 2; same results) */
    int32_t matrix_free;       /* where those paired launches and the FP64 outer row sums take
                                  each row's weights (same results in every form): 0 the nine
                                  stored values; 1 (default) formed from the slot's cell
                                  coefficients (8 bytes per node) with the assembly's
                                  expressions; 2 the paired launches gather them from the
                                  distinct weights (one per cell, edge and node) stored four per
                                  node, interleaved, built with the hierarchy; the FP64 outer
                                  forms them as 1 */
    int32_t first_launch_from_history; /* 1: a honei_hip_batch_step that finds no slot running
                                  (the first launch of a batch) runs as many outer iterations
                                  as the last such batch needed until its last slot retired
                                  (its largest outer iteration count + 2, at most
                                  iters_per_launch; with graphs captured when that batch's
                                  retirement is polled), so a repeated workload finishes in one
                                  launch without trailing iterations over empty lists; later
                                  launches unchanged, results identical.  0 (default): every
                                  launch runs iters_per_launch */
} honei_hip_batch_config;

typedef struct {
    double initial;            /* ||b - A x0|| */
    double norm;               /* latest ||b - A x|| (FP64) */
    double rel;                /* norm / initial */
    int32_t state;             /* 0 idle, 1 running, 2 admitted (first check pending), 3 loading */
    int32_t iters;             /* outer iterations applied */
    int32_t status;            /* 0 converged, 5 not converged, -1 pending */
    int32_t level;             /* finest level of the slot */
} honei_hip_slot_info;

int honei_hip_batch_config_default(honei_hip_batch_config *cfg);
/* levels[c], counts[c] for c < nclass; budget_bytes 0 = no limit, else the
 * pool (matrices, hierarchies, vectors) must fit: HONEI_HIP_ERROR_BUDGET.   */
int honei_hip_batch_create(void **batch, const int64_t *levels, const int64_t *counts, int32_t nclass,
                           const honei_hip_batch_config *cfg, uint64_t budget_bytes);
int honei_hip_batch_destroy(void *batch);
int honei_hip_batch_info(void *batch, int64_t *nslots, uint64_t *device_bytes, int64_t *launches);
/* Global slot index of the j-th slot of class `cls` (caller's class order). */
int honei_hip_batch_slot(void *batch, int32_t cls, int64_t j, int64_t *slot);
/* Assemble slot `slot` (idle, result already taken) from coeff_dev
 * ((side-1)^2 doubles) and b_dev (side^2 doubles); asynchronous, on the
 * batch's load stream.  The slot joins at the next step after the copy. */
int honei_hip_batch_load(void *batch, int64_t slot, const double *coeff_dev, const double *b_dev);
/* Many slots in one bulk build (a handful of launches for all of them):
 * from device arrays (coeff_dev[i], b_dev[i]) or drawn from seeds
 * (coefficient exp(sigma N(0,1)), b uniform in [-1, 1)).                 */
int honei_hip_batch_load_many(void *batch, int64_t n, const int64_t *slots, const double *const *coeff_dev,
                              const double *const *b_dev);
int honei_hip_batch_load_seeded(void *batch, int64_t n, const int64_t *slots, const uint64_t *coeff_seeds,
                                const uint64_t *rhs_seeds, double sigma);
/* Hand loads that have finished (wait != 0: all loads issued so far, blocking)
 * to the next launch; honei_hip_batch_step does this without waiting.     */
int honei_hip_batch_activate(void *batch, int32_t wait);
int honei_hip_batch_step(void *batch);                          /* one graph launch, asynchronous */
int honei_hip_batch_poll(void *batch, honei_hip_slot_info *info); /* waits; info[nslots] */
int honei_hip_batch_result(void *batch, int64_t slot, double *x_dev);
/* Step until every loaded slot has retired; *launches = graph launches. */
int honei_hip_batch_run(void *batch, int64_t *launches);
/* Launch shape of an inner level (threads per row, block size, slots per
 * thread: one index load serves that many systems) and, when the autotuner
 * chose it, the measured time of one batched Jacobi sweep over its tuning
 * batch (microseconds, else 0) and that batch's slot count.                */
int honei_hip_batch_shape(void *batch, int32_t level, int32_t *threads, int32_t *block, int32_t *group,
                          double *sweep_us, int64_t *tune_slots);
/* Levels that run the paired launches (pair_sweeps, levels of 33^2 nodes and
 * up): the autotuner's measured time of one paired Jacobi0 + defect launch
 * and of one paired two-sweep launch over the same tuning batch
 * (microseconds; 0 when the level is not paired or not tuned).            */
int honei_hip_batch_pair_timing(void *batch, int32_t level, double *pre_us, double *post_us);
/* Measurement (no reference counterpart): one level's launches timed over
 * every slot present on that level -- the launch shape of a batched step at
 * full activity -- with HIP events on the pool's stream, `reps` launches
 * after one warm-up: us[0] paired Jacobi0 + defect, us[1] restriction to
 * level - 1, us[2] prolongation + correction from level - 1, us[3] paired
 * two-sweep launch (0 where the level does not pair; where it pairs, the
 * prolongation + correction runs inside it and us[2] is 0), us[4] the FP64
 * p = z + beta p with q = A p, p.q (one tiled launch) over every slot of
 * every class (CG outer).  Overwrites the
 * pool's vectors and solutions: only between steps, with no slot loading or
 * running, after the results are copied out.  Synchronous.               */
int honei_hip_batch_time_level(void *batch, int32_t level, int32_t reps, double *us);

/* ---- serving loop -------------------------------------------------------------------
 * No reference counterpart.  Continuous batching over a batch pool: requests
 * (Q1 systems, side 2^level + 1, coefficient exp(sigma N(0,1)) per cell and
 * b uniform in [-1, 1), drawn on the device from the request's seeds) arrive
 * at arrival_s seconds of wall clock after the call starts; each waits in a
 * FIFO for a free slot of its size class, is assembled on the load stream
 * while admitted work keeps iterating, and joins the next graph launch.
 * Class sizes counts[c] = 0 are derived from budget_bytes (equal share per
 * class, at most max_slots_per_class); the whole pool must fit the budget
 * (HONEI_HIP_ERROR_BUDGET).  results[r].status: 0 converged, 5 not
 * converged, -2 no class for the request's level.  Latency percentiles are
 * over converged requests, from arrival to the poll that saw completion.   */
typedef struct {
    double arrival_s;
    int32_t level;
    int32_t pad;
    uint64_t coeff_seed, rhs_seed;
    double *x_dev;             /* optional: FP64 solution (side^2) copied here */
    /* optional host payload (pinned memory, honei_hip_host_alloc, for
       asynchronous copies): coefficient ((side-1)^2) and right-hand side
       (side^2) copied to the device at admission instead of drawn from the
       seeds (both or neither); the solution (side^2) copied to x_host, the
       request counting as done when that copy has completed               */
    const double *coeff_host;
    const double *b_host;
    double *x_host;
} honei_hip_request;

typedef struct {
    double admit_s, done_s, rel;
    int32_t iters, status;
    int64_t slot;
} honei_hip_request_result;

typedef struct {
    int64_t completed, not_converged, rejected, launches, max_queue;
    double wall_s, solves_per_s, p50_ms, p99_ms, max_ms, mean_ms;
    uint64_t device_bytes;
} honei_hip_serve_stats;

typedef struct {
    honei_hip_batch_config batch;
    int32_t nclass;
    int32_t max_slots_per_class;
    int64_t levels[16];
    int64_t counts[16];
    uint64_t budget_bytes;
    double sigma;
} honei_hip_serve_config;

int honei_hip_serve_run(const honei_hip_serve_config *cfg, const honei_hip_request *reqs, int64_t nreq,
                        honei_hip_request_result *results, honei_hip_serve_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* HONEI_HIP_H */
