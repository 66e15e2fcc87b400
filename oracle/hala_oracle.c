/*
 * hala_oracle.c -- CPU restatement of HALA's sparse iterative-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for libhala_amd.so.  Loaded by
 * tests, the smoke check and the CPU-baseline leg of the benchmark, never by
 * the product.  Built with -O2 -ffp-contract=off (see oracle/Makefile) so no
 * product/sum is fused; the HIP kernels it checks are built the same way.
 *
 * Citations are file:line in the reference tree (LIBHALA/hala).
 */
#include "hala_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CB 256 /* canonical block: elements per block and width of the final pass */

/*
 * Threads.  The default is one.  With more, only work whose per-element
 * operation order does not depend on the split runs in parallel: rows of a
 * product, canonical 256-element blocks of a dot product, elementwise vector
 * updates, rows of one level of a triangular sweep or of the ILU(0)
 * factorisation (a level's rows only read rows of earlier levels).  Every
 * result is therefore bit-identical for any thread count.
 */
static int or_threads = 1;
void hala_or_set_threads(int t) { or_threads = t < 1 ? 1 : t; }
int hala_or_get_threads(void) { return or_threads; }
#define OR_PAR(n) if (or_threads > 1 && (n) > 4096) num_threads(or_threads)

/* Pairwise halving tree over a[0..w): a[t] += a[t+off] for off = w/2 .. 1. */
static double tree_reduce(double *a, int w) {
    for (int off = w / 2; off >= 1; off >>= 1)
        for (int t = 0; t < off; t++) a[t] = a[t] + a[t + off];
    return a[0];
}

/*
 * Canonical dot product.  HALA's dot/norm2 call ddot_/dnrm2_
 * (blas/hala_blas_1.hpp:105-123, :278-291) and leave the summation order to
 * the linked BLAS.  This build fixes one order, shared with every GPU
 * reduction:
 *   1. partial[b] = tree_reduce(x[256b+t]*y[256b+t], t = 0..255), zero padded;
 *   2. s[t] = 0 + partial[t] + partial[t+256] + ... (sequential), t = 0..255;
 *   3. result = tree_reduce(s, 256).
 */
double hala_or_cdot(int64_t n, const double *x, const double *y) {
    if (n <= 0) return 0.0;
    int64_t nb = (n + CB - 1) / CB;
    double *part = (double *)malloc((size_t)nb * sizeof(double));
    double a[CB];
#pragma omp parallel for schedule(static) OR_PAR(n)
    for (int64_t b = 0; b < nb; b++) {
        double w[CB];
        for (int t = 0; t < CB; t++) {
            int64_t i = b * CB + t;
            w[t] = (i < n) ? x[i] * y[i] : 0.0;
        }
        part[b] = tree_reduce(w, CB);
    }
    for (int t = 0; t < CB; t++) {
        double s = 0.0;
        for (int64_t k = t; k < nb; k += CB) s = s + part[k];
        a[t] = s;
    }
    free(part);
    return tree_reduce(a, CB);
}

double hala_or_cnrm2(int64_t n, const double *x) { return sqrt(hala_or_cdot(n, x, x)); }

/*
 * Team dot used inside one workgroup of the batched solver: thread t of a team
 * of `team` threads sums elements t, t+team, ... sequentially from 0, then the
 * team is reduced with the halving tree.
 */
double hala_or_team_dot(int64_t n, const double *x, const double *y, int team) {
    double *a = (double *)malloc((size_t)team * sizeof(double));
    for (int t = 0; t < team; t++) {
        double s = 0.0;
        for (int64_t i = t; i < n; i += team) s = s + x[i] * y[i];
        a[t] = s;
    }
    double r = tree_reduce(a, team);
    free(a);
    return r;
}

/* BLAS-1 in reference-BLAS element order (blas/hala_blas_1.hpp:210-217, :317-335). */
static void axpy(int n, double a, const double *x, double *y) {
#pragma omp parallel for schedule(static) OR_PAR(n)
    for (int i = 0; i < n; i++) y[i] = y[i] + a * x[i];
}
static void scal(int n, double a, double *x) {
#pragma omp parallel for schedule(static) OR_PAR(n)
    for (int i = 0; i < n; i++) x[i] = a * x[i];
}
static void vcopy(int n, const double *x, double *y) {
#pragma omp parallel for schedule(static) OR_PAR(n)
    for (int i = 0; i < n; i++) y[i] = x[i];
}

/*
 * CSR y = alpha op(A) x + beta y.  sparse/hala_sparse_utils.hpp:104-118.
 *  'N': sum = sum + vals[j]*x[indx[j]] in row order, then y = alpha*sum + beta*y.
 *  'T'/'C' (real): y = beta*y, then y[indx[j]] += (alpha*x[i])*vals[j], rows ascending.
 * beta == 0 does not read y (the rocSPARSE/BLAS convention the GPU engine follows);
 * with a finite y this equals HALA's CPU result except for the sign of a zero.
 */
void hala_or_spmv(char trans, int nrows, int ncols, double alpha, const int *pntr, const int *indx,
                  const double *vals, const double *x, double beta, double *y) {
    if (trans == 'N' || trans == 'n') {
#pragma omp parallel for schedule(static) OR_PAR(nrows)
        for (int i = 0; i < nrows; i++) {
            double sum = 0.0;
            for (int j = pntr[i]; j < pntr[i + 1]; j++) sum = sum + vals[j] * x[indx[j]];
            y[i] = (beta == 0.0) ? alpha * sum + 0.0 : alpha * sum + beta * y[i];
        }
    } else {
        if (beta == 0.0) {
            for (int k = 0; k < ncols; k++) y[k] = 0.0;
        } else {
            scal(ncols, beta, y);
        }
        for (int i = 0; i < nrows; i++)
            for (int j = pntr[i]; j < pntr[i + 1]; j++) y[indx[j]] = y[indx[j]] + (alpha * x[i]) * vals[j];
    }
}

/* FP32 'N' product with FP32 accumulation, same order as hala_or_spmv. */
void hala_or_spmv_f32(int nrows, float alpha, const int *pntr, const int *indx, const float *vals,
                      const float *x, float beta, float *y) {
#pragma omp parallel for schedule(static) OR_PAR(nrows)
    for (int i = 0; i < nrows; i++) {
        float sum = 0.0f;
        for (int j = pntr[i]; j < pntr[i + 1]; j++) sum = sum + vals[j] * x[indx[j]];
        y[i] = (beta == 0.0f) ? alpha * sum + 0.0f : alpha * sum + beta * y[i];
    }
}

/*
 * diag[i] = first position in row i whose column is >= i
 * (sparse/hala_sparse_core.hpp:45-60).  Returns the first row whose diagonal
 * entry is missing, or -1 when every row has it (HALA assumes it, :42).
 */
int hala_or_diag(int n, const int *pntr, const int *indx, int *diag) {
    int bad = -1;
    for (int i = 0; i < n; i++) {
        int d = pntr[i];
        while (d < pntr[i + 1] && indx[d] < i) d++;
        diag[i] = d;
        if ((d == pntr[i + 1] || indx[d] != i) && bad < 0) bad = i;
    }
    return bad;
}

/*
 * Level sets of the triangular solves over the combined LU pattern (no HALA
 * function: rocSPARSE's csrsv analysis builds them internally,
 * gpu/hala_rocm_sparse_triangular.hpp:71-76).
 *   lower: lev[i] = 0 without strictly-lower entries, else 1 + max lev[j], j < i in row i;
 *   upper: mirrored over the strictly-upper entries, rows taken from n-1 down.
 * level_ptr[l] = number of rows with level < l (nlev+1 entries); perm lists
 * the rows sorted by (level, row).  Returns nlev.
 */
int hala_or_levels(int n, const int *pntr, const int *indx, const int *diag, int upper, int *lev,
                   int *level_ptr, int *perm) {
    int maxl = -1;
    if (!upper) {
        for (int i = 0; i < n; i++) {
            int m = 0;
            for (int j = pntr[i]; j < diag[i]; j++)
                if (lev[indx[j]] + 1 > m) m = lev[indx[j]] + 1;
            lev[i] = m;
            if (m > maxl) maxl = m;
        }
    } else {
        for (int i = n - 1; i >= 0; i--) {
            int m = 0;
            for (int j = diag[i] + 1; j < pntr[i + 1]; j++)
                if (lev[indx[j]] + 1 > m) m = lev[indx[j]] + 1;
            lev[i] = m;
            if (m > maxl) maxl = m;
        }
    }
    int nlev = maxl + 1;
    for (int l = 0; l <= nlev; l++) level_ptr[l] = 0;
    for (int i = 0; i < n; i++) level_ptr[lev[i] + 1]++;
    for (int l = 0; l < nlev; l++) level_ptr[l + 1] += level_ptr[l];
    int *cur = (int *)malloc((size_t)(nlev + 1) * sizeof(int));
    memcpy(cur, level_ptr, (size_t)(nlev + 1) * sizeof(int));
    for (int i = 0; i < n; i++) perm[cur[lev[i]]++] = i;
    free(cur);
    return nlev;
}

/*
 * Level schedule of both sweeps over the combined pattern, for the threaded
 * factorisation and applies: rows of one level of L (of U) depend only on
 * rows of earlier levels.
 */
typedef struct {
    int nl, nu;
    int *lptr, *lperm, *uptr, *uperm;
} sched_t;

static void sched_free(sched_t *s) {
    if (!s) return;
    free(s->lptr);
    free(s->lperm);
    free(s->uptr);
    free(s->uperm);
    free(s);
}

static sched_t *sched_build(int n, const int *pntr, const int *indx, const int *diag) {
    if (or_threads <= 1 || n <= 4096) return 0;
    sched_t *s = (sched_t *)calloc(1, sizeof(sched_t));
    int *lev = (int *)malloc((size_t)n * sizeof(int));
    s->lptr = (int *)malloc((size_t)(n + 1) * sizeof(int));
    s->uptr = (int *)malloc((size_t)(n + 1) * sizeof(int));
    s->lperm = (int *)malloc((size_t)n * sizeof(int));
    s->uperm = (int *)malloc((size_t)n * sizeof(int));
    s->nl = hala_or_levels(n, pntr, indx, diag, 0, lev, s->lptr, s->lperm);
    s->nu = hala_or_levels(n, pntr, indx, diag, 1, lev, s->uptr, s->uperm);
    free(lev);
    return s;
}

/* the sweeps of one triangle, level by level (rows of a level in parallel) */
#define OR_ILU_APPLY_SCHED(NAME, R)                                                                    \
    static void NAME(const sched_t *s, const int *pntr, const int *indx, const int *diag, const R *ilu, \
                     R *x) {                                                                           \
        _Pragma("omp parallel num_threads(or_threads)") {                                              \
            for (int l = 1; l < s->nl; l++) {                                                          \
                _Pragma("omp for schedule(static)") for (int q = s->lptr[l]; q < s->lptr[l + 1]; q++) { \
                    const int i = s->lperm[q];                                                         \
                    for (int j = pntr[i]; j < diag[i]; j++) x[i] = x[i] - ilu[j] * x[indx[j]];         \
                }                                                                                      \
            }                                                                                          \
            for (int l = 0; l < s->nu; l++) {                                                          \
                _Pragma("omp for schedule(static)") for (int q = s->uptr[l]; q < s->uptr[l + 1]; q++) { \
                    const int i = s->uperm[q];                                                         \
                    for (int j = diag[i] + 1; j < pntr[i + 1]; j++) x[i] = x[i] - ilu[j] * x[indx[j]]; \
                    x[i] = x[i] / ilu[diag[i]];                                                        \
                }                                                                                      \
            }                                                                                          \
        }                                                                                              \
    }
OR_ILU_APPLY_SCHED(ilu_apply_sched, double)
OR_ILU_APPLY_SCHED(ilu_apply_sched_f32, float)

/* IKJ elimination of row j (see hala_or_ilu0) */
static void ilu0_row(int j, const int *pntr, const int *indx, const int *diag, double *ilu) {
    for (int jc = pntr[j]; jc < diag[j]; jc++) {
        int i = indx[jc];
        ilu[jc] = ilu[jc] / ilu[diag[i]];
        double l = ilu[jc];
        int ik = diag[i] + 1, jk = jc + 1;
        while (ik < pntr[i + 1] && jk < pntr[j + 1]) {
            if (indx[ik] == indx[jk]) {
                ilu[jk] = ilu[jk] - l * ilu[ik];
                ik++;
                jk++;
            } else if (indx[ik] < indx[jk]) {
                ik++;
            } else {
                jk++;
            }
        }
    }
}

/*
 * ILU(0) in IKJ (row) order.  HALA's factorize_ilu_array
 * (sparse/hala_sparse_utils.hpp:229-254) loops KIJ: for each pivot i it
 * updates every later row j holding a_ji.  Entry (j,k) then receives the same
 * updates in the same increasing-i order, each reading the same final row i,
 * so the row-order form below performs the identical operation sequence on
 * every entry at O(nnz * row length) cost instead of O(n^2).
 */
void hala_or_ilu0(int n, const int *pntr, const int *indx, const int *diag, double *ilu) {
    sched_t *s = sched_build(n, pntr, indx, diag);
    if (!s) {
        for (int j = 0; j < n; j++) ilu0_row(j, pntr, indx, diag, ilu);
        return;
    }
#pragma omp parallel num_threads(or_threads)
    for (int l = 1; l < s->nl; l++) {
#pragma omp for schedule(static)
        for (int q = s->lptr[l]; q < s->lptr[l + 1]; q++) ilu0_row(s->lperm[q], pntr, indx, diag, ilu);
    }
    sched_free(s);
}

/* HALA's own KIJ loop order (sparse/hala_sparse_utils.hpp:229-254), O(n^2): small inputs only. */
void hala_or_ilu0_kij(int n, const int *pntr, const int *indx, const int *diag, double *ilu) {
    for (int i = 0; i < n; i++) {
        double u = ilu[diag[i]];
        for (int j = i + 1; j < n; j++) {
            int jc = pntr[j];
            while (jc < pntr[j + 1] && indx[jc] < i) jc++;
            if (jc < pntr[j + 1] && indx[jc] == i) {
                ilu[jc] = ilu[jc] / u;
                double l = ilu[jc];
                int ik = diag[i] + 1, jk = jc + 1;
                while (ik < pntr[i + 1] && jk < pntr[j + 1]) {
                    if (indx[ik] == indx[jk]) {
                        ilu[jk] = ilu[jk] - l * ilu[ik];
                        ik++;
                        jk++;
                    } else if (indx[ik] < indx[jk]) {
                        ik++;
                    } else {
                        jk++;
                    }
                }
            }
        }
    }
}

/*
 * x <- U^-1 L^-1 x over the combined factor (sparse/hala_sparse_utils.hpp:263-275):
 * forward x[i] -= L_ij x[j] term by term in row order (unit L); backward
 * x[i] -= U_ij x[j] term by term, then x[i] /= U_ii.
 */
void hala_or_ilu_apply(int n, const int *pntr, const int *indx, const int *diag, const double *ilu, double *x) {
    for (int i = 1; i < n; i++)
        for (int j = pntr[i]; j < diag[i]; j++) x[i] = x[i] - ilu[j] * x[indx[j]];
    for (int i = n - 1; i >= 0; i--) {
        for (int j = diag[i] + 1; j < pntr[i + 1]; j++) x[i] = x[i] - ilu[j] * x[indx[j]];
        x[i] = x[i] / ilu[diag[i]];
    }
}

void hala_or_ilu_apply_f32(int n, const int *pntr, const int *indx, const int *diag, const float *ilu, float *x) {
    for (int i = 1; i < n; i++)
        for (int j = pntr[i]; j < diag[i]; j++) x[i] = x[i] - ilu[j] * x[indx[j]];
    for (int i = n - 1; i >= 0; i--) {
        for (int j = diag[i] + 1; j < pntr[i + 1]; j++) x[i] = x[i] - ilu[j] * x[indx[j]];
        x[i] = x[i] / ilu[diag[i]];
    }
}

/*
 * x <- (LU)^-1 x with the factor applied by k >= 1 Jacobi sweeps per
 * triangle, in FP32 or FP64: a method of this build for the inner solve of
 * iterative refinement, not a HALA operation.  L (unit): y0 = x,
 *   y(m+1)_i = x_i - l_ij y(m)_j - ...   (j < i, one term at a time in row order);
 * U: z0_i = y_i / u_ii,
 *   z(m+1)_i = (y_i - u_ij z(m)_j - ...) / u_ii   (j > i, row order),
 * with y = y(k).  Every row applies its terms in the exact sweep's order, so
 * with k at least the level count of a triangle minus one the result equals
 * the exact apply of the same precision bit for bit.  w: 3n values of scratch.
 */
#define OR_ILU_JACOBI(NAME, R)                                                                                \
    void NAME(int n, const int *pntr, const int *indx, const int *diag, const R *ilu, int k, R *x, R *w) {      \
        R *ya = w, *yb = w + n, *zb = w + 2 * (size_t)n;                                                    \
        const R *yo = x; /* y0 = x */                                                                       \
        R *yn = ya;                                                                                         \
        for (int m = 0; m < k; m++) {                                                                       \
            _Pragma("omp parallel for schedule(static) if (or_threads > 1 && n > 4096) num_threads(or_threads)") \
            for (int i = 0; i < n; i++) {                                                                   \
                R acc = x[i];                                                                               \
                for (int j = pntr[i]; j < diag[i]; j++) acc = acc - ilu[j] * yo[indx[j]];                   \
                yn[i] = acc;                                                                                \
            }                                                                                               \
            yo = yn;                                                                                        \
            yn = (yn == ya) ? yb : ya;                                                                      \
        }                                                                                                   \
        /* y(k) is in yo; z0 into the free y buffer, then alternate it with zb */                          \
        const R *y = yo;                                                                                    \
        R *zo = (yo == ya) ? yb : ya;                                                                       \
        _Pragma("omp parallel for schedule(static) if (or_threads > 1 && n > 4096) num_threads(or_threads)")     \
        for (int i = 0; i < n; i++) zo[i] = y[i] / ilu[diag[i]];                                            \
        R *zn = zb;                                                                                         \
        for (int m = 0; m < k; m++) {                                                                       \
            _Pragma("omp parallel for schedule(static) if (or_threads > 1 && n > 4096) num_threads(or_threads)") \
            for (int i = 0; i < n; i++) {                                                                   \
                R acc = y[i];                                                                               \
                for (int j = diag[i] + 1; j < pntr[i + 1]; j++) acc = acc - ilu[j] * zo[indx[j]];           \
                zn[i] = acc / ilu[diag[i]];                                                                 \
            }                                                                                               \
            R *t = zo;                                                                                      \
            zo = zn;                                                                                        \
            zn = t;                                                                                         \
        }                                                                                                   \
        memcpy(x, zo, (size_t)n * sizeof(R));                                                               \
    }
OR_ILU_JACOBI(hala_or_ilu_apply_f32_jacobi, float)
OR_ILU_JACOBI(hala_or_ilu_apply_jacobi, double)

/*
 * x <- (LU)^-T x = L^-T U^-T x (the adjoint preconditioner; HALA has no
 * transposed ILU apply).  U^T w = x column by column: w_i = x_i / U_ii, then
 * x_k -= U_ik w_i for k > i; L^T y = w from the last row up: x_k -= L_ik x_i
 * for k < i.  Every x_k therefore receives its updates in increasing i
 * (U^T) and decreasing i (L^T) order, one term at a time.
 */
void hala_or_ilu_apply_t(int n, const int *pntr, const int *indx, const int *diag, const double *ilu, double *x) {
    for (int i = 0; i < n; i++) {
        x[i] = x[i] / ilu[diag[i]];
        for (int j = diag[i] + 1; j < pntr[i + 1]; j++) x[indx[j]] = x[indx[j]] - ilu[j] * x[i];
    }
    for (int i = n - 1; i >= 0; i--)
        for (int j = pntr[i]; j < diag[i]; j++) x[indx[j]] = x[indx[j]] - ilu[j] * x[i];
}

/*
 * Standalone triangular solve op(T) x = alpha b for a CSR matrix holding one
 * triangle, the diagonal being the last (L) or first (U) entry of each row
 * (sparse/hala_sparse_utils.hpp:284-335).  'N' sums the off-diagonal terms
 * first, then subtracts; 'T'/'C' is column oriented.
 */
#define OR_TRSV(NAME, R)                                                                                     \
    void NAME(char uplo, char trans, char diag, int n, R alpha, const int *pntr, const int *indx, const R *vals,  \
              const R *b, R *x) {                                                                             \
        int unit = (diag == 'U' || diag == 'u');                                                              \
        int lower = (uplo == 'L' || uplo == 'l');                                                             \
        if (trans == 'N' || trans == 'n') {                                                                   \
            if (lower) {                                                                                      \
                x[0] = unit ? alpha * b[0] : alpha * b[0] / vals[0];                                          \
                for (int i = 1; i < n; i++) {                                                                 \
                    R s = (R)0;                                                                               \
                    for (int j = pntr[i]; j < pntr[i + 1] - 1; j++) s = s + vals[j] * x[indx[j]];             \
                    x[i] = unit ? (alpha * b[i] - s) : (alpha * b[i] - s) / vals[pntr[i + 1] - 1];            \
                }                                                                                             \
            } else {                                                                                          \
                x[n - 1] = unit ? alpha * b[n - 1] : alpha * b[n - 1] / vals[pntr[n] - 1];                    \
                for (int i = n - 2; i >= 0; i--) {                                                            \
                    R s = (R)0;                                                                               \
                    for (int j = pntr[i] + 1; j < pntr[i + 1]; j++) s = s + vals[j] * x[indx[j]];             \
                    x[i] = unit ? (alpha * b[i] - s) : (alpha * b[i] - s) / vals[pntr[i]];                    \
                }                                                                                             \
            }                                                                                                 \
        } else {                                                                                              \
            for (int i = 0; i < n; i++) x[i] = (R)0 * x[i];                                                   \
            if (lower) {                                                                                      \
                for (int i = n - 1; i >= 0; i--) {                                                            \
                    x[i] = unit ? (alpha * b[i] - x[i]) : (alpha * b[i] - x[i]) / vals[pntr[i + 1] - 1];      \
                    for (int j = pntr[i]; j < pntr[i + 1] - 1; j++) x[indx[j]] = x[indx[j]] + x[i] * vals[j]; \
                }                                                                                             \
            } else {                                                                                          \
                for (int i = 0; i < n; i++) {                                                                 \
                    x[i] = unit ? (alpha * b[i] - x[i]) : (alpha * b[i] - x[i]) / vals[pntr[i]];              \
                    for (int j = pntr[i] + 1; j < pntr[i + 1]; j++) x[indx[j]] = x[indx[j]] + x[i] * vals[j]; \
                }                                                                                             \
            }                                                                                                 \
        }                                                                                                     \
    }
OR_TRSV(hala_or_trsv, double)
OR_TRSV(hala_or_trsv_f32, float)

/* ------------------------------------------------------------------------- */
/* Preconditioners as the solvers apply them                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
    int kind;
    int n;
    const int *pntr, *indx, *diag;
    const double *data;
    sched_t *sched; /* level schedule of the ILU sweeps (threaded runs), or null */
} prec_t;

static void prec_apply(const prec_t *m, const double *r, double *z) {
    int n = m->n;
    switch (m->kind) {
    case HALA_OR_PREC_IDENTITY: /* vcopy */
        vcopy(n, r, z);
        break;
    case HALA_OR_PREC_ILU: /* cpu_ilu::apply: copy then sweep, sparse/hala_sparse_ilu.hpp:105-115 */
        vcopy(n, r, z);
        if (m->sched) ilu_apply_sched(m->sched, m->pntr, m->indx, m->diag, m->data, z);
        else hala_or_ilu_apply(n, m->pntr, m->indx, m->diag, m->data, z);
        break;
    case HALA_OR_PREC_JACOBI: /* vdivide, wax/hala_blas_extensions.hpp:139-151 */
#pragma omp parallel for schedule(static) OR_PAR(n)
        for (int i = 0; i < n; i++) z[i] = r[i] / m->data[i];
        break;
    case HALA_OR_PREC_ILU_T:
        memcpy(z, r, (size_t)n * sizeof(double));
        hala_or_ilu_apply_t(n, m->pntr, m->indx, m->diag, m->data, z);
        break;
    }
}

typedef double (*dot_fn)(int64_t, const double *, const double *, int);
static double dot_canon(int64_t n, const double *x, const double *y, int team) {
    (void)team;
    return hala_or_cdot(n, x, y);
}

/*
 * Preconditioned CG with HALA's solve_cg_core step order
 * (hex/solvers/hala_solvers_cg.hpp:92-156 with the lambdas of :203-226):
 * iterations start at 1 and count operator applications; the loop stops when
 * iterations == max_iter or norm2(r) < tol (absolute, recursively updated r).
 * hist[k] receives norm2(r) after the k-th loop pass.
 */
static int cg_core(int n, const int *pntr, const int *indx, const double *vals, const prec_t *m, const double *b,
                   double *x, double tol, int max_iter, dot_fn dot, int team, double *hist, int max_hist) {
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *p = (double *)malloc((size_t)n * sizeof(double));
    double *ap = (double *)malloc((size_t)n * sizeof(double));
    double *z = (double *)malloc((size_t)n * sizeof(double));
    vcopy(n, b, r);
    hala_or_spmv('N', n, n, 1.0, pntr, indx, vals, x, 0.0, p);
    axpy(n, -1.0, p, r);
    prec_apply(m, r, z);
    vcopy(n, z, p);
    double zr = dot(n, r, z, team);
    int it = 1;
    for (;;) {
        hala_or_spmv('N', n, n, 1.0, pntr, indx, vals, p, 0.0, ap);
        it++;
        double nzr = dot(n, p, ap, team);
        double a = zr / nzr;
        axpy(n, a, p, x);
        axpy(n, -a, ap, r);
        double rn = sqrt(dot(n, r, r, team));
        if (hist && it - 2 < max_hist) hist[it - 2] = rn;
        if (it == max_iter || rn < tol) break;
        prec_apply(m, r, z);
        nzr = dot(n, r, z, team);
        a = nzr / zr;
        scal(n, a, p);
        axpy(n, 1.0, z, p);
        zr = nzr;
    }
    free(r);
    free(p);
    free(ap);
    free(z);
    return it;
}

int hala_or_cg(int n, const int *pntr, const int *indx, const double *vals, int prec, const int *mpntr,
               const int *mindx, const int *mdiag, const double *mdata, const double *b, double *x, double tol,
               int max_iter, double *hist, int max_hist) {
    prec_t m = {prec, n, mpntr, mindx, mdiag, mdata, 0};
    if (prec == HALA_OR_PREC_ILU) m.sched = sched_build(n, mpntr, mindx, mdiag);
    int it = cg_core(n, pntr, indx, vals, &m, b, x, tol, max_iter, dot_canon, 0, hist, max_hist);
    sched_free(m.sched);
    return it;
}

/*
 * Multi-RHS PCG with one shared matrix: HALA's solve_batch_cg
 * (hex/solvers/hala_solvers_cg_batch.hpp:68-110), i.e. solve_cg_core with the
 * batch operations of wax/hala_blas_extensions.hpp.  Column k of B, X is
 * entries [k n, (k+1) n); every scalar becomes a vector over the columns:
 *   zr_k = <r_k, z_k>; alpha_k = zr_k / <p_k, Ap_k> (vdivide);
 *   x_k += (1.0 alpha_k) p_k, r_k += (-1.0 alpha_k) Ap_k (batch_axpy, :51-66);
 *   stop when iterations == max_iter or max_k ||r_k|| < tol (batch_max_norm2,
 *   :354-365: the largest <r_k, r_k>, then sqrt);
 *   beta_k = <r_k, z_k>new / zr_k; p_k = beta_k p_k (batch_scal), p += z (axpy).
 * All columns iterate until the stop.  hist receives max_k ||r_k|| per pass.
 */
int hala_or_cg_batch(int n, const int *pntr, const int *indx, const double *vals, int prec, const int *mpntr,
                     const int *mindx, const int *mdiag, const double *mdata, int nrhs, const double *B, double *X,
                     double tol, int max_iter, double *hist, int max_hist) {
    prec_t m = {prec, n, mpntr, mindx, mdiag, mdata, 0};
    const size_t N = (size_t)n * nrhs;
    double *r = (double *)malloc(N * sizeof(double));
    double *p = (double *)malloc(N * sizeof(double));
    double *ap = (double *)malloc(N * sizeof(double));
    double *z = (double *)malloc(N * sizeof(double));
    double *zr = (double *)malloc((size_t)nrhs * sizeof(double));
    double *a = (double *)malloc((size_t)nrhs * sizeof(double));
    memcpy(r, B, N * sizeof(double));
    for (int k = 0; k < nrhs; k++)
        hala_or_spmv('N', n, n, 1.0, pntr, indx, vals, X + (size_t)k * n, 0.0, p + (size_t)k * n);
    axpy((int)N, -1.0, p, r);
    for (int k = 0; k < nrhs; k++) prec_apply(&m, r + (size_t)k * n, z + (size_t)k * n);
    memcpy(p, z, N * sizeof(double));
    for (int k = 0; k < nrhs; k++) zr[k] = hala_or_cdot(n, r + (size_t)k * n, z + (size_t)k * n);
    int it = 1;
    for (;;) {
        for (int k = 0; k < nrhs; k++)
            hala_or_spmv('N', n, n, 1.0, pntr, indx, vals, p + (size_t)k * n, 0.0, ap + (size_t)k * n);
        it++;
        double mx = 0.0;
        for (int k = 0; k < nrhs; k++) {
            const size_t o = (size_t)k * n;
            a[k] = zr[k] / hala_or_cdot(n, p + o, ap + o);
            axpy(n, 1.0 * a[k], p + o, X + o);
            axpy(n, -1.0 * a[k], ap + o, r + o);
            double d = hala_or_cdot(n, r + o, r + o);
            if (k == 0 || d > mx) mx = d;
        }
        double rn = sqrt(mx);
        if (hist && it - 2 < max_hist) hist[it - 2] = rn;
        if (it == max_iter || rn < tol) break;
        for (int k = 0; k < nrhs; k++) prec_apply(&m, r + (size_t)k * n, z + (size_t)k * n);
        for (int k = 0; k < nrhs; k++) {
            const size_t o = (size_t)k * n;
            double nzr = hala_or_cdot(n, r + o, z + o);
            scal(n, nzr / zr[k], p + o);
            zr[k] = nzr;
        }
        axpy((int)N, 1.0, z, p);
    }
    free(r);
    free(p);
    free(ap);
    free(z);
    free(zr);
    free(a);
    return it;
}

/* Jacobi-PCG of one batched system, reductions in team order. */
int hala_or_cg_team(int n, const int *pntr, const int *indx, const double *vals, const double *dvals,
                    const double *b, double *x, double tol, int max_iter, int team, double *hist, int max_hist) {
    prec_t m = {HALA_OR_PREC_JACOBI, n, 0, 0, 0, dvals, 0};
    return cg_core(n, pntr, indx, vals, &m, b, x, tol, max_iter, hala_or_team_dot, team, hist, max_hist);
}

/* ------------------------------------------------------------------------- */
/* Givens rotations and packed triangular solve (reference BLAS)             */
/* ------------------------------------------------------------------------- */

/* drotg, reference BLAS formulation (blas/hala_blas_1.hpp:361-386 binds drotg_). */
void hala_or_rotg(double *a, double *b, double *c, double *s) {
    double sa = *a, sb = *b;
    double roe = (fabs(sa) > fabs(sb)) ? sa : sb;
    double scale = fabs(sa) + fabs(sb);
    double r, z;
    if (scale == 0.0) {
        *c = 1.0;
        *s = 0.0;
        r = 0.0;
        z = 0.0;
    } else {
        double as = sa / scale, bs = sb / scale;
        r = scale * sqrt(as * as + bs * bs);
        r = (roe >= 0.0) ? r : -r;
        *c = sa / r;
        *s = sb / r;
        z = 1.0;
        if (fabs(sa) > fabs(sb)) z = *s;
        if (fabs(sb) >= fabs(sa) && *c != 0.0) z = 1.0 / *c;
    }
    *a = r;
    *b = z;
}

/* drot on one pair (blas/hala_blas_1.hpp:393-410 binds drot_). */
void hala_or_rot(double *x, double *y, double c, double s) {
    double t = c * *x + s * *y;
    *y = c * *y - s * *x;
    *x = t;
}

/* dtpsv('U','N','N'): packed column-major upper triangle, column sweep from the last column. */
void hala_or_tpsv_un(int n, const double *ap, double *x) {
    int kk = n * (n + 1) / 2 - 1;
    for (int j = n - 1; j >= 0; j--) {
        if (x[j] != 0.0) {
            x[j] = x[j] / ap[kk];
            double t = x[j];
            int k = kk - 1;
            for (int i = j - 1; i >= 0; i--) {
                x[i] = x[i] - t * ap[k];
                k--;
            }
        }
        kk -= j + 1;
    }
}

/* ------------------------------------------------------------------------- */
/* GMRES                                                                      */
/* ------------------------------------------------------------------------- */

/* Operator/preconditioner of one GMRES instance. */
typedef struct {
    int n;
    const int *pntr, *indx, *diag;
    const double *vals;
    const float *vals32; /* non-null: FP32 operator and FP32 ILU */
    const double *ilu;
    const float *ilu32;
    int prec; /* IDENTITY or ILU */
    float *w32a, *w32b;
    int jacobi;  /* FP32 ILU: k > 0 Jacobi sweeps per triangle instead of the exact sweeps */
    float *wj;   /* 3n floats for the Jacobi apply */
    double *wj64; /* 3n doubles for the FP64 Jacobi apply */
    sched_t *sched; /* level schedule of the ILU sweeps (threaded runs), or null */
} gm_op_t;

/* t = alpha*A*x + beta*t in FP64, or the FP32 product converted back. */
static void gm_spmv(gm_op_t *o, double alpha, const double *x, double beta, double *t) {
    if (!o->vals32) {
        hala_or_spmv('N', o->n, o->n, alpha, o->pntr, o->indx, o->vals, x, beta, t);
        return;
    }
#pragma omp parallel for schedule(static) OR_PAR(o->n)
    for (int i = 0; i < o->n; i++) o->w32a[i] = (float)x[i];
    hala_or_spmv_f32(o->n, 1.0f, o->pntr, o->indx, o->vals32, o->w32a, 0.0f, o->w32b);
#pragma omp parallel for schedule(static) OR_PAR(o->n)
    for (int i = 0; i < o->n; i++)
        t[i] = (beta == 0.0) ? alpha * (double)o->w32b[i] + 0.0 : alpha * (double)o->w32b[i] + beta * t[i];
}

static void gm_prec(gm_op_t *o, const double *t, double *r) {
    int n = o->n;
    if (o->prec == HALA_OR_PREC_IDENTITY) {
        vcopy(n, t, r);
    } else if (!o->ilu32) {
        vcopy(n, t, r);
        if (o->jacobi > 0) hala_or_ilu_apply_jacobi(n, o->pntr, o->indx, o->diag, o->ilu, o->jacobi, r, o->wj64);
        else if (o->sched) ilu_apply_sched(o->sched, o->pntr, o->indx, o->diag, o->ilu, r);
        else hala_or_ilu_apply(n, o->pntr, o->indx, o->diag, o->ilu, r);
    } else {
#pragma omp parallel for schedule(static) OR_PAR(n)
        for (int i = 0; i < n; i++) o->w32a[i] = (float)t[i];
        if (o->jacobi > 0)
            hala_or_ilu_apply_f32_jacobi(n, o->pntr, o->indx, o->diag, o->ilu32, o->jacobi, o->w32a, o->wj);
        else if (o->sched)
            ilu_apply_sched_f32(o->sched, o->pntr, o->indx, o->diag, o->ilu32, o->w32a);
        else
            hala_or_ilu_apply_f32(n, o->pntr, o->indx, o->diag, o->ilu32, o->w32a);
#pragma omp parallel for schedule(static) OR_PAR(n)
        for (int i = 0; i < n; i++) r[i] = (double)o->w32a[i];
    }
}

/*
 * Restarted left-preconditioned GMRES with HALA's step order
 * (hex/solvers/hala_solvers_gmres.hpp:125-230): one classical Gram-Schmidt
 * pass through gemv 'T' / 'N' (krylov_project, :67-72), Givens rotations on the
 * host, absolute tolerance on the Givens estimate |s_k Z_k|, and the last
 * rotation of a cycle never applied to Z (only the "still going" branch
 * rotates Z, :210-216).  max_outer counts restart cycles; the return value
 * counts preconditioner applications.
 *   tol_rel > 0: the tolerance is tol_rel times the first preconditioned
 *                residual norm of the call (used by the refinement inner solve);
 *   x_zero: x is known to be zero, so b - A x is b itself (no product).
 */
static int gmres_core(gm_op_t *o, const double *b, double *x, double tol, double tol_rel, int x_zero, int restart,
                      int max_outer, double *hist, int max_hist, int *nhist) {
    int n = o->n;
    double *W = (double *)malloc((size_t)n * (size_t)restart * sizeof(double));
    double *t = (double *)malloc((size_t)n * sizeof(double));
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *H = (double *)malloc((size_t)restart * (restart + 1) / 2 * sizeof(double) + sizeof(double));
    double *S = (double *)malloc((size_t)(restart + 1) * sizeof(double));
    double *C = (double *)malloc((size_t)(restart + 1) * sizeof(double));
    double *Z = (double *)malloc((size_t)(restart + 1) * sizeof(double));
    double *cf = (double *)malloc((size_t)(restart + 1) * sizeof(double));
    double inner_res, outer_res = tol + 1.0;
    int total = 0, outer = 0, nh = nhist ? *nhist : 0;
    int first = 1;
    while (outer_res > tol && outer < max_outer) {
        int nH = 0, nZ = 0;
        vcopy(n, b, t);
        if (!(x_zero && outer == 0)) gm_spmv(o, -1.0, x, 1.0, t);
        gm_prec(o, t, r);
        total++;
        inner_res = hala_or_cnrm2(n, r);
        if (first && tol_rel > 0.0) {
            tol = tol_rel * inner_res;
            first = 0;
        }
        scal(n, 1.0 / inner_res, r);
        Z[nZ++] = inner_res;
        vcopy(n, r, W);
        int k = 0;
        while (inner_res > tol && k < restart) {
            gm_spmv(o, 1.0, r, 0.0, t);
            gm_prec(o, t, r);
            total++;
            for (int j = 0; j <= k; j++) cf[j] = hala_or_cdot(n, W + (size_t)j * n, r);
#pragma omp parallel for schedule(static) OR_PAR(n)
            for (int i = 0; i < n; i++) {
                double yi = r[i];
                for (int j = 0; j <= k; j++) yi = yi + (-cf[j]) * W[(size_t)j * n + i];
                r[i] = yi;
            }
            double nrm = hala_or_cnrm2(n, r);
            scal(n, 1.0 / nrm, r);
            for (int i = 0; i < k; i++) hala_or_rot(&cf[i], &cf[i + 1], C[i], S[i]);
            double c, s, beta = nrm;
            hala_or_rotg(&cf[k], &beta, &c, &s);
            for (int j = 0; j <= k; j++) H[nH++] = cf[j];
            S[k] = s;
            C[k] = c;
            inner_res = fabs(s * Z[nZ - 1]);
            if (hist && nh < max_hist) hist[nh] = inner_res;
            nh++;
            k++;
            if (inner_res > tol && k < restart) {
                vcopy(n, r, W + (size_t)k * n);
                Z[nZ++] = 0.0;
                hala_or_rot(&Z[k - 1], &Z[k], c, s);
            }
        }
        if (nH > 0) {
            hala_or_tpsv_un(nZ, H, Z);
#pragma omp parallel for schedule(static) OR_PAR(n)
            for (int i = 0; i < n; i++) {
                double yi = x[i];
                for (int j = 0; j < nZ; j++) yi = yi + (1.0 * Z[j]) * W[(size_t)j * n + i];
                x[i] = yi;
            }
        }
        outer++;
        outer_res = inner_res;
    }
    if (nhist) *nhist = nh;
    free(W);
    free(t);
    free(r);
    free(H);
    free(S);
    free(C);
    free(Z);
    free(cf);
    return total;
}

int hala_or_gmres(int n, const int *pntr, const int *indx, const double *vals, int prec, const int *diag,
                  const double *ilu, const double *b, double *x, double tol, int restart, int max_outer,
                  double *hist, int max_hist, int *nhist) {
    gm_op_t o = {n, pntr, indx, diag, vals, 0, ilu, 0, prec, 0, 0, 0, 0, 0, 0};
    if (nhist) *nhist = 0;
    if (prec == HALA_OR_PREC_ILU) o.sched = sched_build(n, pntr, indx, diag);
    int total = gmres_core(&o, b, x, tol, 0.0, 0, restart, max_outer, hist, max_hist, nhist);
    sched_free(o.sched);
    return total;
}

/*
 * Mixed-precision iterative refinement (no HALA counterpart): the FP64 outer
 * loop forms r = b - A x with the FP64 product, stops when
 * norm2(r) < rtol*norm2(b), and otherwise solves A d = r with GMRES(restart)
 * whose operator and ILU(0) preconditioner are FP32 (FP32 accumulation) while
 * its vectors, Gram-Schmidt, Givens and dot products stay FP64; the inner
 * solve stops at inner_rtol times its first preconditioned residual norm
 * (at most inner_cycles restarts), then x += d.
 * hist receives the outer relative residuals; returns the total number of
 * inner preconditioner applications, *nhist the number of outer residuals.
 */
int hala_or_gmres_ir(int n, const int *pntr, const int *indx, const double *vals, const float *vals32,
                     const int *diag, const double *ilu, const float *ilu32, const double *b, double *x,
                     double rtol, double inner_rtol, int restart, int max_refine, double *hist, int max_hist,
                     int *nhist, int jacobi) {
    const int inner_cycles = 20;
    float *w32a = (float *)malloc((size_t)n * sizeof(float));
    float *w32b = (float *)malloc((size_t)n * sizeof(float));
    float *wj = (float *)malloc((size_t)3 * n * sizeof(float));
    double *wj64 = (double *)malloc((size_t)3 * n * sizeof(double));
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *d = (double *)malloc((size_t)n * sizeof(double));
    gm_op_t o = {n, pntr, indx, diag, vals, vals32, ilu, vals32 ? ilu32 : 0, HALA_OR_PREC_ILU, w32a, w32b,
                 jacobi, wj, wj64, sched_build(n, pntr, indx, diag)};
    double bn = hala_or_cnrm2(n, b);
    int total = 0, nh = 0;
    for (int it = 0; it <= max_refine; it++) {
        vcopy(n, b, r);
        hala_or_spmv('N', n, n, -1.0, pntr, indx, vals, x, 1.0, r);
        double rel = hala_or_cnrm2(n, r) / bn;
        if (hist && nh < max_hist) hist[nh] = rel;
        nh++;
        if (rel < rtol || it == max_refine) break;
        memset(d, 0, (size_t)n * sizeof(double));
        total += gmres_core(&o, r, d, 0.0, inner_rtol, 1, restart, inner_cycles, 0, 0, 0);
        axpy(n, 1.0, d, x);
    }
    if (nhist) *nhist = nh;
    sched_free(o.sched);
    free(w32a);
    free(w32b);
    free(wj);
    free(wj64);
    free(r);
    free(d);
    return total;
}

/* ------------------------------------------------------------------------- */
/* Adjoint pieces                                                             */
/* ------------------------------------------------------------------------- */

/* dL/dA_ij = -lambda_i x_j on every stored entry. */
void hala_or_grad_vals(int n, const int *pntr, const int *indx, const double *lam, const double *x, double *gvals) {
    for (int i = 0; i < n; i++)
        for (int j = pntr[i]; j < pntr[i + 1]; j++) gvals[j] = -(lam[i] * x[indx[j]]);
}

/* Chain rule onto stencil coefficients: g_c = sum_e ccoef[e] * gvals[cnnz[e]], e in cptr[c]..cptr[c+1]. */
void hala_or_grad_coef(int ncoef, const int *cptr, const int *cnnz, const double *ccoef, const double *gvals,
                       double *gcoef) {
    for (int c = 0; c < ncoef; c++) {
        double s = 0.0;
        for (int e = cptr[c]; e < cptr[c + 1]; e++) s = s + ccoef[e] * gvals[cnnz[e]];
        gcoef[c] = s;
    }
}

/* CSR transpose; each transposed row lists its entries by increasing original row. */
void hala_or_transpose(int nrows, int ncols, const int *pntr, const int *indx, const double *vals, int *tpntr,
                       int *tindx, double *tvals) {
    for (int k = 0; k <= ncols; k++) tpntr[k] = 0;
    for (int j = 0; j < pntr[nrows]; j++) tpntr[indx[j] + 1]++;
    for (int k = 0; k < ncols; k++) tpntr[k + 1] += tpntr[k];
    int *cur = (int *)malloc((size_t)(ncols + 1) * sizeof(int));
    memcpy(cur, tpntr, (size_t)(ncols + 1) * sizeof(int));
    for (int i = 0; i < nrows; i++)
        for (int j = pntr[i]; j < pntr[i + 1]; j++) {
            int q = cur[indx[j]]++;
            tindx[q] = i;
            tvals[q] = vals[j];
        }
    free(cur);
}

/* ---- ingestion ----------------------------------------------------------- */
static void coo_merge_sort(int64_t *a, int64_t *tmp, int64_t n, const int *rows, const int *cols) {
    /* bottom-up stable merge sort of entry positions by (row, column) */
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                const int64_t ea = a[i], eb = a[j];
                const int later = rows[eb] < rows[ea] || (rows[eb] == rows[ea] && cols[eb] < cols[ea]);
                tmp[k++] = later ? a[j++] : a[i++];
            }
            while (i < mid) tmp[k++] = a[i++];
            while (j < hi) tmp[k++] = a[j++];
        }
        memcpy(a, tmp, (size_t)n * sizeof(int64_t));
    }
}

int64_t hala_or_coo_to_csr(int nrows, int ncols, int64_t nent, const int *rows, const int *cols, const double *vals,
                           int add_diag, int *pntr, int *indx, double *vals_out) {
    for (int64_t e = 0; e < nent; e++)
        if (rows[e] < 0 || rows[e] >= nrows || cols[e] < 0 || cols[e] >= ncols) return -1 - e;
    int64_t *a = (int64_t *)malloc((size_t)(nent > 0 ? nent : 1) * sizeof(int64_t));
    int64_t *tmp = (int64_t *)malloc((size_t)(nent > 0 ? nent : 1) * sizeof(int64_t));
    for (int64_t e = 0; e < nent; e++) a[e] = e;
    coo_merge_sort(a, tmp, nent, rows, cols);
    int64_t q = 0, o = 0;
    pntr[0] = 0;
    for (int r = 0; r < nrows; r++) {
        int need = add_diag && r < ncols;
        while (q < nent && rows[a[q]] == r) {
            const int c = cols[a[q]];
            if (need && c > r) {
                if (indx) {
                    indx[o] = r;
                    vals_out[o] = 0.0;
                }
                o++;
                need = 0;
            }
            double s = indx ? vals[a[q]] : 0.0;
            for (q++; q < nent && rows[a[q]] == r && cols[a[q]] == c; q++)
                if (indx) s = s + vals[a[q]];
            if (c == r) need = 0;
            if (indx) {
                indx[o] = c;
                vals_out[o] = s;
            }
            o++;
        }
        if (need) {
            if (indx) {
                indx[o] = r;
                vals_out[o] = 0.0;
            }
            o++;
        }
        pntr[r + 1] = (int)o;
    }
    free(a);
    free(tmp);
    return o;
}

/* ---- complex scalars ------------------------------------------------------ */
/* Canonical sum (hala_or_cdot's order) of a term array, in the term type. */
#define HALA_OR_CSUM(R, NAME)                                                 \
    static R NAME(int64_t n, const R *t) {                                    \
        if (n <= 0) return (R)0;                                              \
        int64_t nb = (n + CB - 1) / CB;                                       \
        R *part = (R *)malloc((size_t)nb * sizeof(R));                        \
        R a[CB];                                                              \
        for (int64_t b = 0; b < nb; b++) {                                    \
            for (int q = 0; q < CB; q++) {                                    \
                int64_t i = b * CB + q;                                       \
                a[q] = (i < n) ? t[i] : (R)0;                                 \
            }                                                                 \
            for (int off = CB / 2; off >= 1; off >>= 1)                       \
                for (int q = 0; q < off; q++) a[q] = a[q] + a[q + off];       \
            part[b] = a[0];                                                   \
        }                                                                     \
        for (int q = 0; q < CB; q++) {                                        \
            R s = (R)0;                                                       \
            for (int64_t k = q; k < nb; k += CB) s = s + part[k];             \
            a[q] = s;                                                         \
        }                                                                     \
        free(part);                                                           \
        for (int off = CB / 2; off >= 1; off >>= 1)                           \
            for (int q = 0; q < off; q++) a[q] = a[q] + a[q + off];           \
        return a[0];                                                          \
    }
HALA_OR_CSUM(double, csum_z)
HALA_OR_CSUM(float, csum_c)

#define HALA_OR_COMPLEX(SUF, R, CSUM, SQRT)                                                                          \
    void hala_or_zdot_##SUF(int64_t n, const R *x, const R *y, int conj, R *res) {                                   \
        R *tr = (R *)malloc((size_t)(n > 0 ? n : 1) * sizeof(R)), *ti = (R *)malloc((size_t)(n > 0 ? n : 1) * sizeof(R)); \
        for (int64_t i = 0; i < n; i++) {                                                                            \
            const R xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];                              \
            tr[i] = conj ? xr * yr + xi * yi : xr * yr - xi * yi;                                                    \
            ti[i] = conj ? xr * yi - xi * yr : xr * yi + xi * yr;                                                    \
        }                                                                                                            \
        res[0] = CSUM(n, tr);                                                                                        \
        res[1] = CSUM(n, ti);                                                                                        \
        free(tr);                                                                                                    \
        free(ti);                                                                                                    \
    }                                                                                                                \
    R hala_or_znrm2_##SUF(int64_t n, const R *x) {                                                                   \
        R *t = (R *)malloc((size_t)(n > 0 ? n : 1) * sizeof(R));                                                     \
        for (int64_t i = 0; i < n; i++) t[i] = x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1];                    \
        R r = SQRT(CSUM(n, t));                                                                                      \
        free(t);                                                                                                     \
        return r;                                                                                                    \
    }                                                                                                                \
    /* sparse/hala_sparse_utils.hpp:104-118 element order, complex arithmetic */                                     \
    void hala_or_zspmv_##SUF(char trans, int nrows, int ncols, R are, R aim, const int *pntr, const int *indx,       \
                             const R *vals, const R *x, R bre, R bim, R *y) {                                        \
        const int beta0 = bre == (R)0 && bim == (R)0;                                                                \
        if (trans == 'N' || trans == 'n') {                                                                          \
            for (int i = 0; i < nrows; i++) {                                                                        \
                R sr = (R)0, si = (R)0;                                                                              \
                for (int j = pntr[i]; j < pntr[i + 1]; j++) {                                                        \
                    const R vr = vals[2 * j], vi = vals[2 * j + 1], xr = x[2 * indx[j]], xi = x[2 * indx[j] + 1];     \
                    sr = sr + (vr * xr - vi * xi);                                                                   \
                    si = si + (vr * xi + vi * xr);                                                                   \
                }                                                                                                    \
                R orr = are * sr - aim * si, oi = are * si + aim * sr;                                               \
                if (!beta0) {                                                                                        \
                    const R yr = y[2 * i], yi = y[2 * i + 1];                                                        \
                    orr = orr + (bre * yr - bim * yi);                                                               \
                    oi = oi + (bre * yi + bim * yr);                                                                 \
                }                                                                                                    \
                y[2 * i] = orr;                                                                                      \
                y[2 * i + 1] = oi;                                                                                   \
            }                                                                                                        \
            return;                                                                                                  \
        }                                                                                                            \
        const int cj = trans == 'C' || trans == 'c';                                                                 \
        for (int k = 0; k < ncols; k++) {                                                                            \
            if (beta0) {                                                                                             \
                y[2 * k] = (R)0;                                                                                     \
                y[2 * k + 1] = (R)0;                                                                                 \
            } else {                                                                                                 \
                const R yr = y[2 * k], yi = y[2 * k + 1];                                                            \
                y[2 * k] = bre * yr - bim * yi;                                                                      \
                y[2 * k + 1] = bre * yi + bim * yr;                                                                  \
            }                                                                                                        \
        }                                                                                                            \
        for (int i = 0; i < nrows; i++) {                                                                            \
            const R axr = are * x[2 * i] - aim * x[2 * i + 1], axi = are * x[2 * i + 1] + aim * x[2 * i];             \
            for (int j = pntr[i]; j < pntr[i + 1]; j++) {                                                            \
                const R vr = vals[2 * j], vi = cj ? -vals[2 * j + 1] : vals[2 * j + 1];                              \
                const int k = indx[j];                                                                               \
                y[2 * k] = y[2 * k] + (axr * vr - axi * vi);                                                         \
                y[2 * k + 1] = y[2 * k + 1] + (axr * vi + axi * vr);                                                 \
            }                                                                                                        \
        }                                                                                                            \
    }                                                                                                                \
    /* op(T) x = alpha b for one stored triangle (sparse/hala_sparse_utils.hpp:284-335): 'N' sums the row's      \
       off-diagonal terms v x first; 'T'/'C' is column oriented, x_i * v added to x_k in the sweep's order;        \
       'C' conjugates the values and the diagonal.  Complex product and quotient as documented above. */           \
    void hala_or_ztrsv_##SUF(char uplo, char trans, char diag, int n, R are, R aim, const int *pntr,              \
                             const int *indx, const R *vals, const R *b, R *x) {                                     \
        const int unit = diag == 'U' || diag == 'u', lower = uplo == 'L' || uplo == 'l';                             \
        const int cj = trans == 'C' || trans == 'c';                                                                 \
        if (trans == 'N' || trans == 'n') {                                                                          \
            for (int t = 0; t < n; t++) {                                                                            \
                const int i = lower ? t : n - 1 - t;                                                                 \
                const int js = lower ? pntr[i] : pntr[i] + 1, je = lower ? pntr[i + 1] - 1 : pntr[i + 1];            \
                R sr = (R)0, si = (R)0;                                                                              \
                for (int j = js; j < je; j++) {                                                                      \
                    const R vr = vals[2 * j], vi = vals[2 * j + 1], xr = x[2 * indx[j]], xi = x[2 * indx[j] + 1];     \
                    sr = sr + (vr * xr - vi * xi);                                                                   \
                    si = si + (vr * xi + vi * xr);                                                                   \
                }                                                                                                    \
                R yr = (are * b[2 * i] - aim * b[2 * i + 1]) - sr, yi = (are * b[2 * i + 1] + aim * b[2 * i]) - si;  \
                if (!unit) {                                                                                         \
                    const int d = lower ? pntr[i + 1] - 1 : pntr[i];                                                 \
                    const R dr = vals[2 * d], di = vals[2 * d + 1], q = dr * dr + di * di;                           \
                    const R tr = (yr * dr + yi * di) / q, ti = (yi * dr - yr * di) / q;                              \
                    yr = tr;                                                                                         \
                    yi = ti;                                                                                         \
                }                                                                                                    \
                x[2 * i] = yr;                                                                                       \
                x[2 * i + 1] = yi;                                                                                   \
            }                                                                                                        \
            return;                                                                                                  \
        }                                                                                                            \
        for (int i = 0; i < 2 * n; i++) x[i] = (R)0 * x[i];                                                         \
        for (int t = 0; t < n; t++) {                                                                                \
            const int i = lower ? n - 1 - t : t;                                                                     \
            const int d = lower ? pntr[i + 1] - 1 : pntr[i];                                                         \
            R yr = (are * b[2 * i] - aim * b[2 * i + 1]) - x[2 * i];                                                 \
            R yi = (are * b[2 * i + 1] + aim * b[2 * i]) - x[2 * i + 1];                                             \
            if (!unit) {                                                                                             \
                const R dr = vals[2 * d], di = cj ? -vals[2 * d + 1] : vals[2 * d + 1], q = dr * dr + di * di;       \
                const R tr = (yr * dr + yi * di) / q, ti = (yi * dr - yr * di) / q;                                  \
                yr = tr;                                                                                             \
                yi = ti;                                                                                             \
            }                                                                                                        \
            x[2 * i] = yr;                                                                                           \
            x[2 * i + 1] = yi;                                                                                       \
            const int js = lower ? pntr[i] : pntr[i] + 1, je = lower ? pntr[i + 1] - 1 : pntr[i + 1];                \
            for (int j = js; j < je; j++) {                                                                          \
                const R vr = vals[2 * j], vi = cj ? -vals[2 * j + 1] : vals[2 * j + 1];                              \
                const int k = indx[j];                                                                               \
                x[2 * k] = x[2 * k] + (yr * vr - yi * vi);                                                           \
                x[2 * k + 1] = x[2 * k + 1] + (yr * vi + yi * vr);                                                   \
            }                                                                                                        \
        }                                                                                                            \
    }                                                                                                                \
    /* IKJ ILU(0) (sparse/hala_sparse_utils.hpp:229-254 operation order per entry), in place */                      \
    void hala_or_zilu0_##SUF(int n, const int *pntr, const int *indx, const int *diag, R *f) {                       \
        for (int j = 0; j < n; j++) {                                                                                \
            const int je = pntr[j + 1];                                                                              \
            for (int jc = pntr[j]; jc < diag[j]; jc++) {                                                             \
                const int i = indx[jc];                                                                              \
                const R ar = f[2 * jc], ai = f[2 * jc + 1], br = f[2 * diag[i]], bi = f[2 * diag[i] + 1];             \
                const R d = br * br + bi * bi;                                                                       \
                const R lr = (ar * br + ai * bi) / d, li = (ai * br - ar * bi) / d;                                  \
                f[2 * jc] = lr;                                                                                      \
                f[2 * jc + 1] = li;                                                                                  \
                int ik = diag[i] + 1, jk = jc + 1;                                                                   \
                const int ie = pntr[i + 1];                                                                          \
                while (ik < ie && jk < je) {                                                                         \
                    if (indx[ik] == indx[jk]) {                                                                      \
                        const R ur = f[2 * ik], ui = f[2 * ik + 1];                                                  \
                        f[2 * jk] = f[2 * jk] - (lr * ur - li * ui);                                                 \
                        f[2 * jk + 1] = f[2 * jk + 1] - (lr * ui + li * ur);                                         \
                        ik++;                                                                                        \
                        jk++;                                                                                        \
                    } else if (indx[ik] < indx[jk]) {                                                                \
                        ik++;                                                                                        \
                    } else {                                                                                         \
                        jk++;                                                                                        \
                    }                                                                                                \
                }                                                                                                    \
            }                                                                                                        \
        }                                                                                                            \
    }                                                                                                                \
    /* (LU)^-1 x: unit-L forward sweep then U backward sweep (sparse/hala_sparse_utils.hpp:263-275) */              \
    void hala_or_zilu_apply_##SUF(int n, const int *pntr, const int *indx, const int *diag, const R *f, const R *x,  \
                                  R *r) {                                                                            \
        R *y = (R *)malloc((size_t)(2 * (int64_t)n > 0 ? 2 * (int64_t)n : 1) * sizeof(R));                           \
        for (int i = 0; i < n; i++) {                                                                                \
            R sr = x[2 * i], si = x[2 * i + 1];                                                                      \
            for (int j = pntr[i]; j < diag[i]; j++) {                                                                \
                const R vr = f[2 * j], vi = f[2 * j + 1], yr = y[2 * indx[j]], yi = y[2 * indx[j] + 1];              \
                sr = sr - (vr * yr - vi * yi);                                                                       \
                si = si - (vr * yi + vi * yr);                                                                       \
            }                                                                                                        \
            y[2 * i] = sr;                                                                                           \
            y[2 * i + 1] = si;                                                                                       \
        }                                                                                                            \
        for (int i = n - 1; i >= 0; i--) {                                                                           \
            R sr = y[2 * i], si = y[2 * i + 1];                                                                      \
            for (int j = diag[i] + 1; j < pntr[i + 1]; j++) {                                                        \
                const R vr = f[2 * j], vi = f[2 * j + 1], zr = r[2 * indx[j]], zi = r[2 * indx[j] + 1];              \
                sr = sr - (vr * zr - vi * zi);                                                                       \
                si = si - (vr * zi + vi * zr);                                                                       \
            }                                                                                                        \
            const R br = f[2 * diag[i]], bi = f[2 * diag[i] + 1], d = br * br + bi * bi;                             \
            r[2 * i] = (sr * br + si * bi) / d;                                                                      \
            r[2 * i + 1] = (si * br - sr * bi) / d;                                                                  \
        }                                                                                                            \
        free(y);                                                                                                     \
    }
HALA_OR_COMPLEX(z, double, csum_z, sqrt)
HALA_OR_COMPLEX(c, float, csum_c, sqrtf)
