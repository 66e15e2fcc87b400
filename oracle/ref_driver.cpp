// ref_driver.cpp -- builds oracle/_ref/libhala_ref.so from HALA's own CPU
// sparse kernels, compiled where they lie in the reference tree.
//
// TEST INFRASTRUCTURE ONLY: it pins the C restatement in hala_oracle.c against
// the reference's code on the same inputs.  Only the BLAS-free kernels are
// exposed (everything else in HALA's CPU engine needs an external BLAS).
#include "hala_sparse_utils.hpp"
#include "hala_sparse_core.hpp"

#include <vector>

extern "C" {

// hala::get_diagonal_index, sparse/hala_sparse_core.hpp:45-60
void hala_ref_diag(int n, const int *pntr, const int *indx, int *diag) {
    std::vector<int> p(pntr, pntr + n + 1), ix(indx, indx + pntr[n]), d;
    hala::get_diagonal_index(p, ix, d);
    for (int i = 0; i < n; i++) diag[i] = d[i];
}

// hala::factorize_ilu_array (KIJ), sparse/hala_sparse_utils.hpp:229-254
void hala_ref_ilu0(int n, const int *pntr, const int *indx, const int *diag, double *ilu) {
    hala::factorize_ilu_array<double>(n, pntr, indx, diag, ilu);
}

// hala::apply_ilu_array, sparse/hala_sparse_utils.hpp:263-275
void hala_ref_ilu_apply(int n, const int *pntr, const int *indx, const int *diag, const double *ilu, double *x) {
    hala::apply_ilu_array<double>(n, pntr, indx, diag, ilu, x);
}

// hala::sparse_gemv_array<T,'N'>, sparse/hala_sparse_utils.hpp:104-111
void hala_ref_spmv_n(int nrows, int ncols, double alpha, const int *pntr, const int *indx, const double *vals,
                     const double *x, double beta, double *y) {
    hala::sparse_gemv_array<double, 'N'>(nrows, ncols, alpha, pntr, indx, vals, x, beta, y);
}

void hala_ref_spmv_n_f32(int nrows, int ncols, float alpha, const int *pntr, const int *indx, const float *vals,
                         const float *x, float beta, float *y) {
    hala::sparse_gemv_array<float, 'N'>(nrows, ncols, alpha, pntr, indx, vals, x, beta, y);
}

// hala::apply_ilu_array in single precision
void hala_ref_ilu_apply_f32(int n, const int *pntr, const int *indx, const int *diag, const float *ilu, float *x) {
    hala::apply_ilu_array<float>(n, pntr, indx, diag, ilu, x);
}
}
