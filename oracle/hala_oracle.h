/*
 * hala_oracle.h -- CPU restatement of HALA's sparse iterative-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * gfx950 kernels in libhala_amd.so.  Only tests, the smoke check and the CPU
 * baseline leg of bench.py load it; the product never links or calls it.
 *
 * Every routine restates one reference routine (file:line in the reference
 * tree, LIBHALA/hala) with the same per-element operation order, so that a
 * kernel that follows the same order is bit-identical to it:
 *   - products and sums are never fused (the library is built with
 *     -ffp-contract=off, and so are the HIP kernels it checks);
 *   - every reduction that HALA delegates to an external BLAS (ddot_, dnrm2_,
 *     dgemv_ 'T') uses the canonical blocked order hala_or_cdot() defines,
 *     because the reference leaves that order to the linked BLAS.
 *
 * Sparse matrices are zero-based CSR with int32 pntr[n+1], indx[nnz], sorted
 * columns and the diagonal present in every row (sparse/hala_sparse_core.hpp:42).
 */
#ifndef HALA_ORACLE_H
#define HALA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- canonical reductions ------------------------------------------------ */
/* Threads of the oracle (default 1); results are bit-identical for any count. */
void hala_or_set_threads(int t);
int hala_or_get_threads(void);

double hala_or_cdot(int64_t n, const double *x, const double *y);
double hala_or_cnrm2(int64_t n, const double *x);
double hala_or_team_dot(int64_t n, const double *x, const double *y, int team);

/* ---- CSR products -------------------------------------------------------- */
void hala_or_spmv(char trans, int nrows, int ncols, double alpha, const int *pntr, const int *indx,
                  const double *vals, const double *x, double beta, double *y);
void hala_or_spmv_f32(int nrows, float alpha, const int *pntr, const int *indx, const float *vals,
                      const float *x, float beta, float *y);

/* ---- complex scalars (interleaved real, imaginary) ------------------------- */
/* The reference instantiates its routines for complex<float/double>
 * (tests/solvers_tests.cpp:14-40).  Products and quotients:
 *   a*b = (ar br - ai bi, ar bi + ai br);  a/b = ((ar br + ai bi)/d, (ai br - ar bi)/d),
 *   d = br br + bi bi.  dot = conjugated (blas/hala_blas_1.hpp dot vs dotu);
 * dotc/dotu/nrm2 sum per-element terms in the canonical order (hala_or_cdot). */
void hala_or_zdot_z(int64_t n, const double *x, const double *y, int conj, double *res);
void hala_or_zdot_c(int64_t n, const float *x, const float *y, int conj, float *res);
double hala_or_znrm2_z(int64_t n, const double *x);
float hala_or_znrm2_c(int64_t n, const float *x);
void hala_or_zspmv_z(char trans, int nrows, int ncols, double are, double aim, const int *pntr, const int *indx,
                     const double *vals, const double *x, double bre, double bim, double *y);
void hala_or_zspmv_c(char trans, int nrows, int ncols, float are, float aim, const int *pntr, const int *indx,
                     const float *vals, const float *x, float bre, float bim, float *y);
void hala_or_zilu0_z(int n, const int *pntr, const int *indx, const int *diag, double *f);
void hala_or_zilu0_c(int n, const int *pntr, const int *indx, const int *diag, float *f);
void hala_or_zilu_apply_z(int n, const int *pntr, const int *indx, const int *diag, const double *f, const double *x,
                          double *r);
void hala_or_zilu_apply_c(int n, const int *pntr, const int *indx, const int *diag, const float *f, const float *x,
                          float *r);

/* ---- ingestion ----------------------------------------------------------- */
/* COO to CSR (no reference counterpart; the reference takes canonical CSR,
 * sparse/hala_sparse_core.hpp:42): entries ordered by (row, column, input
 * position) with a stable merge sort, duplicates summed left to right in input
 * order, add_diag inserts 0 for a missing diagonal.  indx == NULL: pntr only.
 * Returns nnz, or -1 - e for the first entry e out of range. */
int64_t hala_or_coo_to_csr(int nrows, int ncols, int64_t nent, const int *rows, const int *cols, const double *vals,
                           int add_diag, int *pntr, int *indx, double *vals_out);

/* ---- symbolic analysis --------------------------------------------------- */
int hala_or_diag(int n, const int *pntr, const int *indx, int *diag);
int hala_or_levels(int n, const int *pntr, const int *indx, const int *diag, int upper,
                   int *lev, int *level_ptr, int *perm);

/* ---- ILU(0) --------------------------------------------------------------- */
void hala_or_ilu0(int n, const int *pntr, const int *indx, const int *diag, double *ilu);
void hala_or_ilu0_kij(int n, const int *pntr, const int *indx, const int *diag, double *ilu);
void hala_or_ilu_apply(int n, const int *pntr, const int *indx, const int *diag, const double *ilu, double *x);
void hala_or_ilu_apply_f32(int n, const int *pntr, const int *indx, const int *diag, const float *ilu, float *x);
void hala_or_ilu_apply_f32_jacobi(int n, const int *pntr, const int *indx, const int *diag, const float *ilu, int k,
                                  float *x, float *w);
void hala_or_ilu_apply_jacobi(int n, const int *pntr, const int *indx, const int *diag, const double *ilu, int k,
                              double *x, double *w);
void hala_or_ilu_apply_t(int n, const int *pntr, const int *indx, const int *diag, const double *ilu, double *x);

/* ---- standalone triangular solve ---------------------------------------- */
void hala_or_trsv(char uplo, char trans, char diag, int n, double alpha, const int *pntr, const int *indx,
                  const double *vals, const double *b, double *x);
void hala_or_trsv_f32(char uplo, char trans, char diag, int n, float alpha, const int *pntr, const int *indx,
                      const float *vals, const float *b, float *x);
void hala_or_ztrsv_z(char uplo, char trans, char diag, int n, double are, double aim, const int *pntr,
                     const int *indx, const double *vals, const double *b, double *x);
void hala_or_ztrsv_c(char uplo, char trans, char diag, int n, float are, float aim, const int *pntr, const int *indx,
                     const float *vals, const float *b, float *x);

/* ---- Krylov solvers ------------------------------------------------------ */
/* preconditioner kinds: z = r, z = (LU)^-1 r, z = r ./ d, z = (LU)^-T r */
enum { HALA_OR_PREC_IDENTITY = 0, HALA_OR_PREC_ILU = 1, HALA_OR_PREC_JACOBI = 2, HALA_OR_PREC_ILU_T = 3 };

int hala_or_cg(int n, const int *pntr, const int *indx, const double *vals, int prec,
               const int *mpntr, const int *mindx, const int *mdiag, const double *mdata,
               const double *b, double *x, double tol, int max_iter, double *hist, int max_hist);
int hala_or_cg_batch(int n, const int *pntr, const int *indx, const double *vals, int prec, const int *mpntr,
                     const int *mindx, const int *mdiag, const double *mdata, int nrhs, const double *B, double *X,
                     double tol, int max_iter, double *hist, int max_hist);
int hala_or_cg_team(int n, const int *pntr, const int *indx, const double *vals, const double *dvals,
                    const double *b, double *x, double tol, int max_iter, int team, double *hist, int max_hist);
int hala_or_gmres(int n, const int *pntr, const int *indx, const double *vals, int prec, const int *diag,
                  const double *ilu, const double *b, double *x, double tol, int restart, int max_outer,
                  double *hist, int max_hist, int *nhist);
/* vals32 != NULL: FP32 product and FP32 ILU (ilu32) inside; NULL: FP64 product and ILU (ilu) inside.
   jacobi > 0: the inner ILU (FP32 or FP64) is applied by that many Jacobi sweeps per triangle. */
int hala_or_gmres_ir(int n, const int *pntr, const int *indx, const double *vals, const float *vals32,
                     const int *diag, const double *ilu, const float *ilu32, const double *b, double *x,
                     double rtol, double inner_rtol, int restart, int max_refine, double *hist, int max_hist,
                     int *nhist, int jacobi);

/* ---- Givens / packed triangle (BLAS reference semantics) ----------------- */
void hala_or_rotg(double *a, double *b, double *c, double *s);
void hala_or_rot(double *x, double *y, double c, double s);
void hala_or_tpsv_un(int n, const double *ap, double *x);

/* ---- adjoint -------------------------------------------------------------- */
void hala_or_grad_vals(int n, const int *pntr, const int *indx, const double *lam, const double *x, double *gvals);
void hala_or_grad_coef(int ncoef, const int *cptr, const int *cnnz, const double *ccoef, const double *gvals,
                       double *gcoef);
void hala_or_transpose(int nrows, int ncols, const int *pntr, const int *indx, const double *vals,
                       int *tpntr, int *tindx, double *tvals);

#ifdef __cplusplus
}
#endif
#endif
