// ilujac.hip -- the FP32 ILU(0) factor applied by Jacobi sweeps, for the FP32
// inner solve of iterative refinement (a method of this build, not a HALA
// operation; selected with the context option ilu32_jacobi = k > 0).
//
// (LU)^-1 x with k sweeps per triangle:
//   L (unit): y0 = x,  y(m+1)_i = x_i - l_ij y(m)_j - ...          (j < i)
//   U:        z0 = y / d,  z(m+1)_i = (y_i - u_ij z(m)_j - ...) / d_i  (j > i)
// with y = y(k).  Each row subtracts its terms one at a time in row order, the
// order of the exact sweep, so k at least a triangle's level count minus one
// reproduces the exact FP32 apply bit for bit; fewer sweeps give an
// approximate preconditioner whose error the FP64 outer loop removes.  Every
// sweep is an independent pass over one packed triangle (FP32 values, 256-row
// LDS-staged blocks as in the SpMV): bound by HBM bandwidth instead of the
// dependency latency of the exact sweeps.
#include "ilu.hpp"
#include "util.hpp"

namespace hala_amd {

namespace {

__global__ void jac_count(int n, const int *__restrict__ pntr, const int *__restrict__ diag, int *__restrict__ cl,
                          int *__restrict__ cu) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cl[i] = diag[i] - pntr[i];
    cu[i] = pntr[i + 1] - diag[i] - 1;
}

// strict-L and strict-U parts of every row, columns and FP32 values in row order; FP32 diagonal
__global__ void jac_fill(int n, const int *__restrict__ pntr, const int *__restrict__ indx,
                         const int *__restrict__ diag, const double *__restrict__ f, const int *__restrict__ lptr,
                         const int *__restrict__ uptr, int *__restrict__ lidx, float *__restrict__ lval,
                         int *__restrict__ uidx, float *__restrict__ uval, float *__restrict__ dval) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = pntr[i], d = diag[i], e = pntr[i + 1];
    int o = lptr[i];
    for (int j = a; j < d; j++, o++) {
        lidx[o] = indx[j];
        lval[o] = (float)f[j];
    }
    o = uptr[i];
    for (int j = d + 1; j < e; j++, o++) {
        uidx[o] = indx[j];
        uval[o] = (float)f[j];
    }
    dval[i] = (float)f[d];
}

// One sweep over a packed triangle: xn_i = (b_i - v_ij xo_j - ...) [/ d_i].
// ZINIT: also z0_i = xn_i / d_i (the last L sweep starts the U sweeps; d is U's diagonal).
template <class TB, class TX, class TO, bool UNIT, bool ZINIT, bool NT>
__global__ __launch_bounds__(256) void jac_sweep(int n, const int *__restrict__ ptr, const int *__restrict__ idx,
                                                 const float *__restrict__ val, const float *__restrict__ d,
                                                 const TB *__restrict__ b, const TX *__restrict__ xo,
                                                 TO *__restrict__ xn, float *__restrict__ z0, const int *gate) {
    if (gate && *gate) return;
    __shared__ float prod[spmv_tile<float>::value];
    const int i = blockIdx.x * SPMV_ROWS + threadIdx.x;
    const float init = i < n ? (float)b[i] : 0.0f;
    float acc = csr_block_rowsum<float, TX, float, true, NT>(n, ptr, idx, val, xo, prod, blockIdx.x, init);
    if (i >= n) return;
    if (!UNIT) acc = acc / d[i];
    xn[i] = (TO)acc;
    if (ZINIT) z0[i] = acc / d[i];
}

template <class TB, class TX, class TO, bool UNIT, bool ZINIT>
void sweep(hala_ctx c, bool nt, int n, const int *ptr, const int *idx, const float *val, const float *d, const TB *b,
           const TX *xo, TO *xn, float *z0, const int *gate) {
    const int g = (n + SPMV_ROWS - 1) / SPMV_ROWS;
    if (nt) jac_sweep<TB, TX, TO, UNIT, ZINIT, true><<<g, 256, 0, c->stream>>>(n, ptr, idx, val, d, b, xo, xn, z0, gate);
    else jac_sweep<TB, TX, TO, UNIT, ZINIT, false><<<g, 256, 0, c->stream>>>(n, ptr, idx, val, d, b, xo, xn, z0, gate);
    HALA_CHECK_LAUNCH();
}

}  // namespace

void ilu_jacobi_load(hala_ilu M) {
    hala_ctx c = M->ctx;
    const int n = M->n, g = (n + 255) / 256;
    if (!M->jac_pattern) {
        dbuf cl, cu;
        cl.reserve((size_t)n * sizeof(int));
        cu.reserve((size_t)n * sizeof(int));
        M->jl_ptr.reserve((size_t)(n + 1) * sizeof(int));
        M->ju_ptr.reserve((size_t)(n + 1) * sizeof(int));
        jac_count<<<g, 256, 0, c->stream>>>(n, M->pntr, M->diag.as<int>(), cl.as<int>(), cu.as<int>());
        HALA_CHECK_LAUNCH();
        exclusive_scan_i32(c, cl.as<int>(), M->jl_ptr.as<int>(), n);
        exclusive_scan_i32(c, cu.as<int>(), M->ju_ptr.as<int>(), n);
        const int nl = fetch_i32(c, M->jl_ptr.as<int>() + n), nu = fetch_i32(c, M->ju_ptr.as<int>() + n);
        M->jl_idx.reserve((size_t)std::max(nl, 1) * sizeof(int));
        M->jl_val.reserve((size_t)std::max(nl, 1) * sizeof(float));
        M->ju_idx.reserve((size_t)std::max(nu, 1) * sizeof(int));
        M->ju_val.reserve((size_t)std::max(nu, 1) * sizeof(float));
        M->jd_val.reserve((size_t)n * sizeof(float));
        M->jwork.reserve((size_t)4 * n * sizeof(float));
        M->jac_pattern = true;
    }
    jac_fill<<<g, 256, 0, c->stream>>>(n, M->pntr, M->indx, M->diag.as<int>(), M->factor.as<double>(),
                                       M->jl_ptr.as<int>(), M->ju_ptr.as<int>(), M->jl_idx.as<int>(),
                                       M->jl_val.as<float>(), M->ju_idx.as<int>(), M->ju_val.as<float>(),
                                       M->jd_val.as<float>());
    HALA_CHECK_LAUNCH();
    M->jac_values = true;
}

void ilu_apply32x64_jacobi(hala_ilu M, const double *x, double *r, const int *gate, int k) {
    hala_ctx c = M->ctx;
    HALA_REQUIRE(k >= 1, HALA_ERR_ARG, "Jacobi ILU apply needs at least one sweep");
    if (!M->jac_values) ilu_jacobi_load(M);
    const int n = M->n;
    float *w = M->jwork.as<float>();
    float *ybuf[2] = {w, w + n}, *zbuf[2] = {w + 2 * (size_t)n, w + 3 * (size_t)n};
    // cached loads: the k sweeps of a triangle re-read it back to back, from the MALL
    // (non-temporal streams measured 20 % slower here)
    const bool nt = false;
    const int *lp = M->jl_ptr.as<int>(), *li = M->jl_idx.as<int>();
    const int *up = M->ju_ptr.as<int>(), *ui = M->ju_idx.as<int>();
    const float *lv = M->jl_val.as<float>(), *uv = M->ju_val.as<float>(), *dv = M->jd_val.as<float>();
    // L sweeps: y(m+1) = x - (L - I) y(m), y0 = x; the last one also writes z0 = y / d
    for (int m = 0; m < k; m++) {
        float *yn = ybuf[m & 1];
        const bool last = m == k - 1;
        if (m == 0) {
            if (last) sweep<double, double, float, true, true>(c, nt, n, lp, li, lv, dv, x, x, yn, zbuf[0], gate);
            else sweep<double, double, float, true, false>(c, nt, n, lp, li, lv, dv, x, x, yn, nullptr, gate);
        } else {
            const float *yo = ybuf[(m - 1) & 1];
            if (last) sweep<double, float, float, true, true>(c, nt, n, lp, li, lv, dv, x, yo, yn, zbuf[0], gate);
            else sweep<double, float, float, true, false>(c, nt, n, lp, li, lv, dv, x, yo, yn, nullptr, gate);
        }
    }
    // U sweeps: z(m+1) = (y - (U - D) z(m)) / d; the last writes the FP64 result
    const float *y = ybuf[(k - 1) & 1];
    for (int m = 0; m < k; m++) {
        const float *zo = zbuf[m & 1];
        if (m == k - 1) sweep<float, float, double, false, false>(c, nt, n, up, ui, uv, dv, y, zo, r, nullptr, gate);
        else sweep<float, float, float, false, false>(c, nt, n, up, ui, uv, dv, y, zo, zbuf[(m + 1) & 1], nullptr, gate);
    }
}

}  // namespace hala_amd
