// adjoint.hip -- gradient pieces of the adjoint solve (no reference
// counterpart: HALA has no differentiation).
//
// For A(k) x = f and a loss L(x), the adjoint lambda solves A^T lambda = dL/dx;
// then dL/dA_ij = -lambda_i x_j on every stored entry, and a coefficient c
// feeding entries e with weights w_e gets dL/dc = sum_e w_e dL/dA_e.
#include "spmv.hpp"

namespace hala_amd {

__global__ void grad_vals_kernel(int n, const int *__restrict__ pntr, const int *__restrict__ indx,
                                 const double *__restrict__ lam, const double *__restrict__ x,
                                 double *__restrict__ g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double li = lam[i];
    for (int j = pntr[i]; j < pntr[i + 1]; j++) g[j] = -(li * x[indx[j]]);
}

// out[c] = sum_{e in ptr[c]..ptr[c+1]} w[e] * src[idx[e]], sequential from 0
__global__ void gather_sum_kernel(int m, const int *__restrict__ ptr, const int *__restrict__ idx,
                                  const double *__restrict__ w, const double *__restrict__ src,
                                  double *__restrict__ out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    double s = 0.0;
    for (int e = ptr[c]; e < ptr[c + 1]; e++) s = s + w[e] * src[idx[e]];
    out[c] = s;
}

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

int hala_grad_vals_f64(hala_csr A, const double *lam, const double *x, double *gvals) {
    return guard([&] {
        HALA_REQUIRE(A && lam && x && gvals, HALA_ERR_ARG, "hala_grad_vals_f64: null argument");
        set_device(A->ctx);
        if (A->nrows == 0) return;
        grad_vals_kernel<<<(A->nrows + 255) / 256, 256, 0, A->ctx->stream>>>(A->nrows, A->pntr, A->indx, lam, x,
                                                                             gvals);
        HALA_CHECK_LAUNCH();
    });
}

int hala_grad_coef_f64(hala_ctx c, int ncoef, const int *cptr, const int *cnnz, const double *ccoef,
                       const double *gvals, double *gcoef) {
    return guard([&] {
        HALA_REQUIRE(c && ncoef >= 0 && cptr && cnnz && ccoef && gvals && gcoef, HALA_ERR_ARG,
                     "hala_grad_coef_f64: bad argument");
        set_device(c);
        if (ncoef == 0) return;
        gather_sum_kernel<<<(ncoef + 255) / 256, 256, 0, c->stream>>>(ncoef, cptr, cnnz, ccoef, gvals, gcoef);
        HALA_CHECK_LAUNCH();
    });
}

int hala_assemble_f64(hala_ctx c, int64_t nnz, const int *aptr, const int *acoef_idx, const double *aw,
                      const double *coef, double *vals) {
    return guard([&] {
        HALA_REQUIRE(c && nnz >= 0 && nnz < (1ll << 31) && aptr && acoef_idx && aw && coef && vals, HALA_ERR_ARG,
                     "hala_assemble_f64: bad argument");
        set_device(c);
        if (nnz == 0) return;
        gather_sum_kernel<<<(int)((nnz + 255) / 256), 256, 0, c->stream>>>((int)nnz, aptr, acoef_idx, aw, coef, vals);
        HALA_CHECK_LAUNCH();
    });
}

}  // extern "C"
