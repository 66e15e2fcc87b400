/*
 * hala_amd.h -- C ABI of libhala_amd.so, the gfx950 (MI355X) sparse iterative
 * solve engine behind the hala:: templates of hala.hpp and hala_solvers.hpp.
 *
 * Plain C: device pointers, sizes and status codes only.  Every entry point
 * returns HALA_OK (0) or a negative status; hala_last_error() gives the
 * message of the last failure on the calling thread.  The C++ templates turn
 * a non-zero status into std::runtime_error, as HALA does for library errors
 * (gpu/hala_rocm_common.hpp:88-109 in the reference).
 *
 * Data layout: CSR with zero-based int32 pntr[nrows+1], indx[nnz], columns
 * sorted within each row and the diagonal present in square matrices
 * (sparse/hala_sparse_core.hpp:42).  Matrix and vector arguments are device
 * pointers on the context's device unless a name says "host".
 *
 * Each declaration cites the reference interface it replaces (file:line in
 * LIBHALA/hala).
 */
#ifndef HALA_AMD_H
#define HALA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status -------------------------------------------------------------- */
#define HALA_OK 0
#define HALA_ERR_ARG (-1)       /* invalid argument / size */
#define HALA_ERR_HIP (-2)       /* HIP runtime error */
#define HALA_ERR_ALLOC (-3)     /* device or host allocation failed */
#define HALA_ERR_PATTERN (-4)   /* unsorted columns, column out of range, missing diagonal */
#define HALA_ERR_PIVOT (-5)     /* zero or non-finite pivot in ILU(0) */
#define HALA_ERR_NOCONV (-6)    /* a device dependency wait gave up (never expected) */
#define HALA_ERR_INTERNAL (-7)

const char *hala_last_error(void);
int hala_version(int *major, int *minor);

typedef struct hala_ctx_s *hala_ctx;
typedef struct hala_csr_s *hala_csr;
typedef struct hala_ilu_s *hala_ilu;
typedef struct hala_tri_s *hala_tri;
typedef struct hala_batch_s *hala_batch;

/* ---- engine: device, stream --------------------------------------------- */
/* gpu_engine(int deviceid), device(), set_active_device():
 * gpu/hala_gpu_engine.hpp:167-192, :242-245; gpu_device_count():
 * gpu/hala_rocm_common.hpp:120-125.  stream == NULL uses the null stream. */
int hala_device_count(int *count);
int hala_ctx_create(int device, void *stream, hala_ctx *ctx);
int hala_ctx_destroy(hala_ctx ctx);
int hala_ctx_set_stream(hala_ctx ctx, void *stream);
int hala_ctx_sync(hala_ctx ctx);
int hala_ctx_device(hala_ctx ctx, int *device);
/* Options (no reference counterpart; the reference's rocSPARSE calls pick
 * their own).  Launch shapes: "tri_waves_per_cu", "tri_pool_waves",
 * "tri_sleep", "tri_streams", "tri_col_cols", "tri_col_split",
 * "cg_batch_max", "spmv_blocks_per_cu" (the fused SpMV + dot of CG: 0, the
 * default, launches a workgroup per 256-row block, k > 0 at most k per CU over
 * a grid-stride loop).  Paths: "tri_col" (1: column-blocked
 * sweeps where the plan fits, 0: level pool), "tri_stencil" (1: pencil sweeps
 * for stencil patterns, read when the ILU object is created), "cg_fuse" (1:
 * the sweeps' permutations fused into the PCG vector kernels), "cg_xcd" (1:
 * with cg_fuse, the r update and its scatter into the L sweep's input in
 * XCD-aware block order; 0: dispatch order).  Methods:
 * "ilu32_jacobi", "ilu64_jacobi" (k > 0: k Jacobi sweeps per triangle instead
 * of the exact apply, inside GMRES and the standalone applies).  Errors:
 * "sync_checks" (1: standalone sweeps wait for the device and report a
 * dependency wait that gave up; 0: the report is deferred to the next
 * synchronising call).  Unknown names return HALA_ERR_ARG. */
int hala_ctx_set_option(hala_ctx ctx, const char *name, int value);
int hala_ctx_get_option(hala_ctx ctx, const char *name, int *value);

/* ---- memory: gpu_vector, gpu_copy_n, fill -------------------------------- */
/* gpu_allocate/gpu_free/gpu_copy_n: gpu/hala_rocm_common.hpp:233-297;
 * gpu_vector::fill: gpu/hala_gpu_vector.hpp:147-155 (one kernel here instead
 * of log2(n) synchronous copies).  kind: 0 host->device, 1 device->host,
 * 2 device->device. */
int hala_malloc(hala_ctx ctx, size_t bytes, void **ptr);
int hala_free(hala_ctx ctx, void *ptr);
int hala_memcpy(hala_ctx ctx, void *dst, const void *src, size_t bytes, int kind);
int hala_fill_f64(hala_ctx ctx, double *x, int64_t n, double value);
int hala_fill_f32(hala_ctx ctx, float *x, int64_t n, float value);
int hala_fill_i32(hala_ctx ctx, int *x, int64_t n, int value);
int hala_convert_f64_f32(hala_ctx ctx, int64_t n, const double *x, float *y);
int hala_convert_f32_f64(hala_ctx ctx, int64_t n, const float *x, double *y);

/* ---- BLAS-1 on the engine ------------------------------------------------ */
/* dot: gpu/hala_gpu_blas1.hpp:179-198; norm2: :102-121; axpy: :205-222;
 * scal: :229-245; vcopy: :49-77.  Reductions follow the canonical order
 * (256-element blocks, halving tree, sequential pass over block partials, a
 * final 256-wide tree); result_host is a host pointer. */
int hala_dot_f64(hala_ctx ctx, int64_t n, const double *x, const double *y, double *result_host);
int hala_nrm2_f64(hala_ctx ctx, int64_t n, const double *x, double *result_host);
int hala_axpy_f64(hala_ctx ctx, int64_t n, double alpha, const double *x, double *y);
int hala_scal_f64(hala_ctx ctx, int64_t n, double alpha, double *x);
int hala_copy_f64(hala_ctx ctx, int64_t n, const double *x, double *y);
int hala_dot_f32(hala_ctx ctx, int64_t n, const float *x, const float *y, float *result_host);
int hala_nrm2_f32(hala_ctx ctx, int64_t n, const float *x, float *result_host);
int hala_axpy_f32(hala_ctx ctx, int64_t n, float alpha, const float *x, float *y);
int hala_scal_f32(hala_ctx ctx, int64_t n, float alpha, float *x);
int hala_copy_f32(hala_ctx ctx, int64_t n, const float *x, float *y);
/* vdivide z = x ./ y (the Jacobi apply): wax/hala_blas_extensions.hpp:139-151, :267-275 */
int hala_vdivide_f64(hala_ctx ctx, int64_t n, const double *x, const double *y, double *z);
/* dense gemv on a column-major m x n matrix: gpu/hala_gpu_blas2.hpp:40-60 */
int hala_gemv_f64(hala_ctx ctx, char trans, int m, int n, double alpha, const double *A, int lda,
                  const double *x, double beta, double *y);

/* ---- CSR matrix and products --------------------------------------------- */
/* make_sparse_matrix(gpu_engine, ...) / gpu_sparse_matrix:
 * gpu/hala_rocm_sparse_general.hpp:74-185.  The handle aliases pntr/indx
 * (non-owning, as the reference) and records row-length statistics that pick
 * the product kernel.  hala_csr_validate checks sorted in-range columns
 * (and, if require_diag, a diagonal entry in every row). */
typedef struct {
    int nrows, ncols;
    int64_t nnz;
    int max_row, min_row;
    double mean_row;
    int kernel; /* 0: LDS-staged block-of-rows, 1: wavefront per row */
    int index_bytes; /* 2: the block-of-rows products stream each column as an
                        int16 offset from its 256-row block's first row (every
                        offset fits: banded patterns); 4: the caller's indx */
} hala_csr_info_t;
int hala_csr_create(hala_ctx ctx, int nrows, int ncols, int64_t nnz, const int *pntr, const int *indx,
                    hala_csr *csr);
int hala_csr_destroy(hala_csr csr);
int hala_csr_info(hala_csr csr, hala_csr_info_t *info);
int hala_csr_set_kernel(hala_csr csr, int kernel);
int hala_csr_validate(hala_csr csr, int require_diag, int *bad_row);

/* gpu_sparse_matrix::gemv -> rocsparse_?csrmv: gpu/hala_rocm_sparse_general.hpp:115-128;
 * free sparse_gemv(gpu_engine, ...): :188-199.  y = alpha op(A) x + beta y,
 * trans 'N' or 'T' ('C' == 'T' for real types).  'N' sums each row in CSR
 * order (sparse/hala_sparse_utils.hpp:104-111); 'T' adds the terms of
 * y_k in increasing source-row order after scaling by beta (:112-117).
 * beta == 0 does not read y. */
int hala_spmv_f64(hala_csr A, char trans, double alpha, const double *vals, const double *x, double beta,
                  double *y);
int hala_spmv_f32(hala_csr A, char trans, float alpha, const float *vals, const float *x, float beta, float *y);
/* FP32 values, FP32 products and accumulation, FP64 vectors:
 * y = alpha * double(sum_f32(vals32 * float(x))) + beta y. */
int hala_spmv_f32x64(hala_csr A, double alpha, const float *vals, const double *x, double beta, double *y);
/* sparse_gemm / gpu_sparse_matrix::gemm (gpu/hala_rocm_sparse_general.hpp:131-161):
 * C = alpha op(A) op(B) + beta C, B b_rows x b_cols and C column-major with
 * leading dimensions ldb, ldc; op(B) is B ('N'), B^T ('T') or B^H ('C').
 * Column by column in the CPU engine's order (sparse/hala_sparse_utils.hpp:127-167),
 * each column one hala_spmv product (bit-identical to it); beta == 0 does not
 * read C.  Complex variants: hala_spmm_z / hala_spmm_c (alpha, beta as re, im). */
int hala_spmm_f64(hala_csr A, char transa, char transb, int b_rows, int b_cols, double alpha, const double *vals,
                  const double *B, int ldb, double beta, double *C, int ldc);
int hala_spmm_f32(hala_csr A, char transa, char transb, int b_rows, int b_cols, float alpha, const float *vals,
                  const float *B, int ldb, float beta, float *C, int ldc);

/* ---- ILU(0) ------------------------------------------------------------- */
/* gpu_ilu<T> and make_ilu(gpu_engine, ...): gpu/hala_gpu_ilu.hpp:45-195.
 * create: symbolic analysis on the device (diagonal positions, lower and
 * upper level sets, level-ordered permutations, the chunked solve layout);
 * factor: numeric ILU(0) of vals, in place on an owned copy (:85-104), level
 * by level in the reference's operation order; apply: r = U^-1 L^-1 x
 * (:150-172) with the reference CPU operation order
 * (sparse/hala_sparse_utils.hpp:263-275); apply_t: r = (LU)^-T x (adjoint).
 * want_f32 keeps an FP32 copy of the factor for the mixed-precision apply. */
typedef struct {
    int n;
    int64_t nnz;
    int nlev_lower, nlev_upper;
    int nchunk_lower, nchunk_upper;
    int64_t slots_lower, slots_upper; /* padded entries of the chunked layout */
} hala_ilu_info_t;
int hala_ilu_create(hala_ctx ctx, int n, int64_t nnz, const int *pntr, const int *indx, hala_ilu *ilu);
int hala_ilu_destroy(hala_ilu ilu);
int hala_ilu_info(hala_ilu ilu, hala_ilu_info_t *info);
/* Layout of the triangular sweeps (no reference counterpart: rocSPARSE's
 * analysis is opaque).  out[8] = {column-blocked sweep in use (0/1), columns,
 * rows per cell (threads per column workgroup), LDS window levels, most
 * external values one cell stages, reason code when not in use, columns per
 * slab of planes (1: contiguous row blocks), rows per plane (0: not split)}. */
int hala_ilu_sweep_info(hala_ilu ilu, int upper, int *out);
/* Whether the sweeps of the ILU(0) apply run on the pencil schedule of a
 * stencil pattern (no reference counterpart).  The pattern is examined when
 * the ILU object is created with the context option "tri_stencil" set to 1
 * (default 0: the general sweeps, no stencil plan).  out[5] = {stencil kind
 * (0: not examined, or not a supported stencil, 1: 2-D 5-point, 2: 2-D
 * 9-point, 3: 3-D 7-point, 4: 3-D 27-point), nx, ny, nz, in use (the option
 * is on now)}. */
int hala_ilu_stencil_info(hala_ilu ilu, int *out);
int hala_ilu_factor_f64(hala_ilu ilu, const double *vals, int want_f32);
int hala_ilu_apply_f64(hala_ilu ilu, const double *x, double *r);
int hala_ilu_apply_f32(hala_ilu ilu, const float *x, float *r);
int hala_ilu_apply_f32x64(hala_ilu ilu, const double *x, double *r);
int hala_ilu_apply_t_f64(hala_ilu ilu, const double *x, double *r);
/* Copies of the symbolic and numeric results (device destinations; sizes:
 * diag n, lev n, level_ptr nlev+1, perm n, factor nnz). */
int hala_ilu_get_symbolic(hala_ilu ilu, int upper, int *diag, int *lev, int *level_ptr, int *perm);
int hala_ilu_get_factor_f64(hala_ilu ilu, double *factor);

/* ---- standalone triangular matrix ---------------------------------------- */
/* make_triangular_matrix(gpu_engine, uplo, diag, ...) / trsv:
 * gpu/hala_rocm_sparse_triangular.hpp:29-115, :185-194, semantics of
 * sparse_trsv_array (sparse/hala_sparse_utils.hpp:284-335): the matrix holds
 * one triangle with the diagonal last ('L') or first ('U') in each row. */
int hala_tri_create(hala_ctx ctx, char uplo, char diag, int n, int64_t nnz, const int *pntr, const int *indx,
                    hala_tri *tri);
int hala_tri_destroy(hala_tri tri);
int hala_trsv_f64(hala_tri tri, char trans, double alpha, const double *vals, const double *b, double *x);
/* The same for float and complex<double> ("z") / complex<float> ("c") values
 * (interleaved real, imaginary), trans N, T or C ('C' conjugates the values):
 * the reference instantiates the triangular matrix for the four scalar types
 * (gpu/hala_rocm_sparse_triangular.hpp:99-115, tests/solvers_tests.cpp:14-40). */
int hala_trsv_f32(hala_tri tri, char trans, float alpha, const float *vals, const float *b, float *x);
int hala_trsv_z(hala_tri tri, char trans, double alpha_re, double alpha_im, const double *vals, const double *b,
                double *x);
int hala_trsv_c(hala_tri tri, char trans, float alpha_re, float alpha_im, const float *vals, const float *b, float *x);
/* sparse_trsm / gpu_triangular_matrix::trsm (gpu/hala_rocm_sparse_triangular.hpp:116-156;
 * semantics of the CPU engine's trsm, sparse/hala_sparse_structs.hpp:278-300):
 * B <- op(T)^-1 (alpha B) column by column.  transb 'N': column i is B[i*ldb ..
 * i*ldb + n); 'T' or 'C': column i is B[i + r*ldb], r < n (the conjugation of B
 * cancels on both sides).  ldb < 0 selects the minimum (n, or nrhs). */
int hala_trsm_f64(hala_tri tri, char transa, char transb, int nrhs, double alpha, const double *vals, double *B,
                  int ldb);
int hala_trsm_f32(hala_tri tri, char transa, char transb, int nrhs, float alpha, const float *vals, float *B, int ldb);
int hala_trsm_z(hala_tri tri, char transa, char transb, int nrhs, double alpha_re, double alpha_im, const double *vals,
                double *B, int ldb);
int hala_trsm_c(hala_tri tri, char transa, char transb, int nrhs, float alpha_re, float alpha_im, const float *vals,
                float *B, int ldb);

/* ---- solvers --------------------------------------------------------------- */
/* solve_cg / solve_cg_ilu on gpu_engine: hex/solvers/hala_solvers_cg.hpp:181-361
 * (loop solve_cg_core :92-156).  prec: 0 identity, 1 ILU(0) (ilu required),
 * 2 Jacobi (pdiag = diagonal values).  Stops when iterations == max_iter or
 * norm2(r) < tol (absolute); *iterations counts operator applications from 1.
 * hist (device, nullable) receives norm2(r) per pass, at most max_hist. */
int hala_cg_f64(hala_csr A, const double *vals, int prec, hala_ilu ilu, const double *pdiag, const double *b,
                double *x, double tol, int max_iter, int *iterations, double *hist, int max_hist);
/* hala_cg_f64 without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results as hala_cg_f64; the work vectors and reduction partials live in a
 * static device array of 64 MiB per GPU, so A may have at most 2^21 rows
 * (larger: HALA_ERR_ARG).  Calls that overlap on one GPU share that array and
 * race; the preconditioner (ilu) is built beforehand, as for hala_cg_f64. */
int hala_cg_f64_noalloc(hala_csr A, const double *vals, int prec, hala_ilu ilu, const double *pdiag,
                        const double *b, double *x, double tol, int max_iter, int *iterations, double *hist,
                        int max_hist);
/* Multi-RHS PCG with one shared matrix, resident on the device: the
   solve_batch_cg / solve_batch_cg_ilu of hex/solvers/hala_solvers_cg_batch.hpp:68-249
   (solve_cg_core with the batch operations of wax/hala_blas_extensions.hpp).
   B, X: n x nrhs column-major (column k at offset k n); X holds the initial
   guess.  prec: 0 identity, 1 ILU (ilu).  Every scalar is per column; all
   columns iterate until iterations == max_iter or max_k ||r_k|| < tol.
   hist (nullable) receives max_k ||r_k|| per pass. */
int hala_cg_batch_f64(hala_csr A, const double *vals, int prec, hala_ilu ilu, int nrhs, const double *B, double *X,
                      double tol, int max_iter, int *iterations, double *hist, int max_hist);
/* Adjoint CG: A^T y = b preconditioned with (LU)^-T. */
int hala_cg_t_f64(hala_csr A, const double *vals, hala_ilu ilu, const double *b, double *x, double tol,
                  int max_iter, int *iterations, double *hist, int max_hist);

/* solve_gmres / solve_gmres_ilu on gpu_engine: hex/solvers/hala_solvers_gmres.hpp:125-267.
 * prec 0 identity, 1 ILU(0).  max_outer counts restart cycles; *total counts
 * preconditioner applications; hist (device) gets the Givens estimates. */
int hala_gmres_f64(hala_csr A, const double *vals, int prec, hala_ilu ilu, const double *b, double *x,
                   double tol, int restart, int max_outer, int *total, double *hist, int max_hist, int *nhist);
/* Mixed precision (new entry point): FP32 ILU(0) and FP32 product inside
 * GMRES(restart), FP64 Arnoldi/Givens/dots and FP64 iterative refinement to
 * norm2(b - A x) < rtol * norm2(b).  ilu must be factored with want_f32.
 * vals32 == NULL runs the same refinement with the FP64 product and ILU inside
 * (the all-FP64 path the mixed one is compared with). */
int hala_gmres_ir_f64(hala_csr A, const double *vals, const float *vals32, hala_ilu ilu, const double *b,
                      double *x, double rtol, double inner_rtol, int restart, int max_refine, int *total,
                      double *hist, int max_hist, int *nhist);

/* ---- batched independent systems ----------------------------------------- */
/* New entry points (HALA's solve_batch_cg, hex/solvers/hala_solvers_cg_batch.hpp:68-249,
 * shares one matrix across right-hand sides; these solve nsys different
 * matrices).  Packed CSR: row_off[nsys+1] system row offsets, pntr[n_total+1]
 * absolute positions, indx local to each system (at most 8128 rows each).
 * Every system runs Jacobi-PCG (z = r ./ dvals) with HALA's CG step order to
 * norm2(r_s) < rtol * norm2(b_s), in one persistent launch replayed from a
 * captured HIP graph; iters (device, nullable) receives each count.
 * hala_batch_tune times the launch shapes and keeps the fastest. */
int hala_batch_create(hala_ctx ctx, int nsys, const int *row_off, const int *pntr, const int *indx,
                      hala_batch *batch);
int hala_batch_destroy(hala_batch batch);
int hala_batch_set_shape(hala_batch batch, int team, int wg_per_cu);
int hala_batch_get_shape(hala_batch batch, int *team, int *wg_per_cu);
int hala_batch_pcg_f64(hala_batch batch, const double *vals, const double *dvals, const double *b, double *x,
                       double rtol, int max_iter, int *iters);
int hala_batch_tune(hala_batch batch, const double *vals, const double *dvals, const double *b, double *x,
                    double rtol, int max_iter, int reps, int *team, int *wg_per_cu, double *ms);

/* ---- adjoint pieces ------------------------------------------------------- */
/* gvals[j] = -lambda[row(j)] * x[indx[j]]; gcoef[c] = sum_e ccoef[e] * gvals[cnnz[e]]. */
int hala_grad_vals_f64(hala_csr A, const double *lam, const double *x, double *gvals);
int hala_grad_coef_f64(hala_ctx ctx, int ncoef, const int *cptr, const int *cnnz, const double *ccoef,
                       const double *gvals, double *gcoef);
/* Values from coefficients: vals[j] = sum_e aw[e] * coef[acoef_idx[e]], e in aptr[j]..aptr[j+1]
 * (sequential from 0), e.g. the stencil entries from per-face conductivities. */
int hala_assemble_f64(hala_ctx ctx, int64_t nnz, const int *aptr, const int *acoef_idx, const double *aw,
                      const double *coef, double *vals);

/* ---- complex scalars: std::complex<double> ("z") and <float> ("c") arrays,
 * real and imaginary parts interleaved.  The reference instantiates its
 * BLAS-1, CSR product and ILU(0) for both (tests/solvers_tests.cpp:14-40,
 * gpu/hala_gpu_blas1.hpp, gpu/hala_rocm_sparse_general.hpp,
 * gpu/hala_gpu_ilu.hpp).  res: two reals (real, imaginary) for the dots, one
 * for nrm2.  dotc conjugates x (HALA's dot on complex vectors), dotu does not.
 * trans 'N', 'T' or 'C'.  The ILU factor reuses the handle's symbolic analysis. */
int hala_dotc_z(hala_ctx ctx, int64_t n, const double *x, const double *y, double *res);
int hala_dotu_z(hala_ctx ctx, int64_t n, const double *x, const double *y, double *res);
int hala_nrm2_z(hala_ctx ctx, int64_t n, const double *x, double *res);
int hala_axpy_z(hala_ctx ctx, int64_t n, double alpha_re, double alpha_im, const double *x, double *y);
int hala_scal_z(hala_ctx ctx, int64_t n, double alpha_re, double alpha_im, double *x);
int hala_spmv_z(hala_csr csr, char trans, double alpha_re, double alpha_im, const double *vals, const double *x,
                double beta_re, double beta_im, double *y);
int hala_ilu_factor_z(hala_ilu ilu, const double *vals);
int hala_ilu_apply_z(hala_ilu ilu, const double *x, double *r);
int hala_dotc_c(hala_ctx ctx, int64_t n, const float *x, const float *y, float *res);
int hala_dotu_c(hala_ctx ctx, int64_t n, const float *x, const float *y, float *res);
int hala_nrm2_c(hala_ctx ctx, int64_t n, const float *x, float *res);
int hala_axpy_c(hala_ctx ctx, int64_t n, float alpha_re, float alpha_im, const float *x, float *y);
int hala_scal_c(hala_ctx ctx, int64_t n, float alpha_re, float alpha_im, float *x);
int hala_spmv_c(hala_csr csr, char trans, float alpha_re, float alpha_im, const float *vals, const float *x,
                float beta_re, float beta_im, float *y);
int hala_spmm_z(hala_csr A, char transa, char transb, int b_rows, int b_cols, double alpha_re, double alpha_im,
                const double *vals, const double *B, int ldb, double beta_re, double beta_im, double *C, int ldc);
int hala_spmm_c(hala_csr A, char transa, char transb, int b_rows, int b_cols, float alpha_re, float alpha_im,
                const float *vals, const float *B, int ldb, float beta_re, float beta_im, float *C, int ldc);
int hala_ilu_factor_c(hala_ilu ilu, const float *vals);
int hala_ilu_apply_c(hala_ilu ilu, const float *x, float *r);

/* ---- ingestion (no reference counterpart: HALA takes CSR arrays that already
 * have ascending unique columns and, for ILU(0), the diagonal,
 * sparse/hala_sparse_core.hpp:42, :56-59) ----------------------------------- */
enum { HALA_MM_REAL = 0, HALA_MM_INTEGER = 1, HALA_MM_PATTERN = 2 };
enum { HALA_MM_GENERAL = 0, HALA_MM_SYMMETRIC = 1, HALA_MM_SKEW = 2 };
/* Matrix Market header (host): sizes, entries stored in the file and an upper
 * bound of the entries after symmetric expansion; field and symmetry codes.
 * Coordinate (real, integer, pattern) and array (real, integer) files;
 * general, symmetric and skew-symmetric. */
int hala_mm_info(const char *path, int *nrows, int *ncols, int64_t *stored, int64_t *expanded, int *field,
                 int *symmetry);
/* Entries into host COO arrays (0-based), in file order; with expand != 0 the
 * mirror (j,i) of an off-diagonal symmetric entry follows it (negated when
 * skew-symmetric).  Array files give every stored position.  *count = entries. */
int hala_mm_read_coo(const char *path, int expand, int64_t capacity, int *rows, int *cols, double *vals,
                     int64_t *count);
/* Host CSR to a "coordinate real general" file, %.17g values (round-trips bits). */
int hala_mm_write_csr(const char *path, int nrows, int ncols, const int *pntr, const int *indx, const double *vals);
/* Device COO to CSR with ascending unique columns: entries ordered by (row,
 * column, position), duplicates summed left to right in input order,
 * add_diag != 0 inserts a zero for every missing diagonal entry.  Call with
 * indx = NULL for pntr (nrows + 1) and *nnz only, then with indx / vals_out of
 * *nnz entries.  Entries out of range give HALA_ERR_PATTERN. */
int hala_coo_to_csr(hala_ctx ctx, int nrows, int ncols, int64_t nent, const int *rows, const int *cols,
                    const double *vals, int add_diag, int *pntr, int *indx, double *vals_out, int64_t *nnz);

/* ---- synthetic problems (host generators, deterministic, splitmix64) ----- */
/* Each fills caller-provided host arrays; call with NULL arrays first to get
 * sizes in *n / *nnz.  Stencils are sorted CSR with Dirichlet rows trimmed. */
int hala_gen_laplace2d(int nx, int ny, double diag, double off, int *n, int64_t *nnz, int *pntr, int *indx,
                       double *vals);
int hala_gen_laplace3d(int nx, int ny, int nz, int *n, int64_t *nnz, int *pntr, int *indx, double *vals);
int hala_gen_discretize_pde(int num_side, int *n, int64_t *nnz, int *pntr, int *indx, double *vals, double *f);
int hala_gen_convdiff27(int m, uint64_t seed, double peclet, int *n, int64_t *nnz, int *pntr, int *indx,
                        double *vals);
int hala_gen_uniform(uint64_t seed, int64_t n, double lo, double hi, double *out);
int hala_gen_batch_laplace(int nsys, uint64_t seed, int side_lo, int side_hi, int *n_total, int64_t *nnz_total,
                           int *sys_rows, int *row_off, int64_t *nnz_off, int *pntr, int *indx, double *vals,
                           double *dvals);
int hala_gen_diffusion3d(int m, uint64_t seed, int *n, int64_t *nnz, int *ncoef, int64_t *nmap, double *kappa,
                         int *pntr, int *indx, double *vals, int *cptr, int *cnnz, double *ccoef);
int hala_gen_gaussian_source(int m, double sigma, double *f);

#ifdef __cplusplus
}
#endif
#endif
