// ilujac.hip -- the ILU(0) factor applied by Jacobi sweeps, for the inner solve
// of iterative refinement (a method of this build, not a HALA operation;
// selected with the context options ilu32_jacobi = k > 0 for the FP32 factor
// and ilu64_jacobi = k > 0 for the FP64 one).
//
// (LU)^-1 x with k sweeps per triangle:
//   L (unit): y0 = x,  y(m+1)_i = x_i - l_ij y(m)_j - ...          (j < i)
//   U:        z0 = y / d,  z(m+1)_i = (y_i - u_ij z(m)_j - ...) / d_i  (j > i)
// with y = y(k).  Each row subtracts its terms one at a time in row order, the
// order of the exact sweep, so k at least a triangle's level count minus one
// reproduces the exact apply bit for bit; fewer sweeps give an approximate
// preconditioner whose error the FP64 outer loop removes.  Every sweep is an
// independent pass over one packed triangle (256-row LDS-staged blocks as in
// the SpMV): bound by HBM bandwidth instead of the dependency latency of the
// exact sweeps.  The FP32 and FP64 variants run the same kernels in their
// value type, so their only difference is the precision.
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

__global__ void jac_fill_idx(int n, const int *__restrict__ pntr, const int *__restrict__ indx,
                             const int *__restrict__ diag, const int *__restrict__ lptr, const int *__restrict__ uptr,
                             int *__restrict__ lidx, int *__restrict__ uidx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = pntr[i], d = diag[i], e = pntr[i + 1];
    int o = lptr[i];
    for (int j = a; j < d; j++, o++) lidx[o] = indx[j];
    o = uptr[i];
    for (int j = d + 1; j < e; j++, o++) uidx[o] = indx[j];
}

// strict-L and strict-U values of every row in row order and the diagonal, in V
template <class V>
__global__ void jac_fill_val(int n, const int *__restrict__ pntr, const int *__restrict__ diag,
                             const double *__restrict__ f, const int *__restrict__ lptr, const int *__restrict__ uptr,
                             V *__restrict__ lval, V *__restrict__ uval, V *__restrict__ dval) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = pntr[i], d = diag[i], e = pntr[i + 1];
    int o = lptr[i];
    for (int j = a; j < d; j++, o++) lval[o] = (V)f[j];
    o = uptr[i];
    for (int j = d + 1; j < e; j++, o++) uval[o] = (V)f[j];
    dval[i] = (V)f[d];
}

// One sweep over a packed triangle: xn_i = (b_i - v_ij xo_j - ...) [/ d_i], in V.
// ZINIT: also z0_i = xn_i / d_i (the last L sweep starts the U sweeps; d is U's diagonal).
template <class V, class TB, class TX, class TO, bool UNIT, bool ZINIT, bool NT, class TI>
__global__ __launch_bounds__(256) void jac_sweep(int n, const int *__restrict__ ptr, const TI *__restrict__ idx,
                                                 const V *__restrict__ val, const V *__restrict__ d,
                                                 const TB *__restrict__ b, const TX *__restrict__ xo,
                                                 TO *__restrict__ xn, V *__restrict__ z0, const int *gate) {
    if (gate && *gate) return;
    __shared__ V prod[spmv_tile<V>::value];
    const int i = blockIdx.x * SPMV_ROWS + threadIdx.x;
    const V init = i < n ? (V)b[i] : V(0);
    V acc = csr_block_rowsum<V, TX, V, true, NT, TI>(n, ptr, idx, val, xo, prod, blockIdx.x, init);
    if (i >= n) return;
    if (!UNIT) acc = acc / d[i];
    xn[i] = (TO)acc;
    if (ZINIT) z0[i] = acc / d[i];
}

// off (nullable): the triangle's columns as int16 block offsets, read instead of idx
template <class V, class TB, class TX, class TO, bool UNIT, bool ZINIT>
void sweep(hala_ctx c, int n, const int *ptr, const int *idx, const short *off, const V *val, const V *d,
           const TB *b, const TX *xo, TO *xn, V *z0, const int *gate) {
    // cached loads: the k sweeps of a triangle re-read it back to back, from the MALL
    // (non-temporal streams measured 20 % slower here)
    const int g = (n + SPMV_ROWS - 1) / SPMV_ROWS;
    if (off)
        jac_sweep<V, TB, TX, TO, UNIT, ZINIT, false, short><<<g, 256, 0, c->stream>>>(n, ptr, off, val, d, b, xo, xn,
                                                                                     z0, gate);
    else
        jac_sweep<V, TB, TX, TO, UNIT, ZINIT, false, int><<<g, 256, 0, c->stream>>>(n, ptr, idx, val, d, b, xo, xn, z0,
                                                                                   gate);
    HALA_CHECK_LAUNCH();
}

// the packed pattern (pointers, columns) of both strict triangles, shared by both precisions
void jac_pattern(hala_ilu M) {
    if (M->jac_pattern) return;
    hala_ctx c = M->ctx;
    const int n = M->n, g = (n + 255) / 256;
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
    M->ju_idx.reserve((size_t)std::max(nu, 1) * sizeof(int));
    jac_fill_idx<<<g, 256, 0, c->stream>>>(n, M->pntr, M->indx, M->diag.as<int>(), M->jl_ptr.as<int>(),
                                           M->ju_ptr.as<int>(), M->jl_idx.as<int>(), M->ju_idx.as<int>());
    HALA_CHECK_LAUNCH();
    // banded triangles: the sweeps stream int16 block offsets instead (6 instead of 8 bytes per FP32 entry)
    if (!build_block_offsets(c, n, nl, M->jl_ptr.as<int>(), M->jl_idx.as<int>(), M->jl_off)) M->jl_off.release();
    if (!build_block_offsets(c, n, nu, M->ju_ptr.as<int>(), M->ju_idx.as<int>(), M->ju_off)) M->ju_off.release();
    M->jac_pattern = true;
}

// values of precision V from the FP64 factor, and four iterates of V
template <class V>
void jac_values(hala_ilu M, dbuf &lv, dbuf &uv, dbuf &dv, dbuf &work) {
    jac_pattern(M);
    hala_ctx c = M->ctx;
    const int n = M->n, g = (n + 255) / 256;
    const int nl = fetch_i32(c, M->jl_ptr.as<int>() + n), nu = fetch_i32(c, M->ju_ptr.as<int>() + n);
    lv.reserve((size_t)std::max(nl, 1) * sizeof(V));
    uv.reserve((size_t)std::max(nu, 1) * sizeof(V));
    dv.reserve((size_t)n * sizeof(V));
    work.reserve((size_t)4 * n * sizeof(V));
    jac_fill_val<V><<<g, 256, 0, c->stream>>>(n, M->pntr, M->diag.as<int>(), M->factor.as<double>(),
                                              M->jl_ptr.as<int>(), M->ju_ptr.as<int>(), lv.as<V>(), uv.as<V>(),
                                              dv.as<V>());
    HALA_CHECK_LAUNCH();
}

// (LU)^-1 x, k sweeps per triangle in precision V, FP64 x and r
template <class V>
void jac_apply(hala_ilu M, const double *x, double *r, const int *gate, int k, const V *lv, const V *uv,
               const V *dv, V *w) {
    hala_ctx c = M->ctx;
    const int n = M->n;
    V *ybuf[2] = {w, w + n}, *zbuf[2] = {w + 2 * (size_t)n, w + 3 * (size_t)n};
    const int *lp = M->jl_ptr.as<int>(), *li = M->jl_idx.as<int>();
    const int *up = M->ju_ptr.as<int>(), *ui = M->ju_idx.as<int>();
    const short *lo = M->jl_off.as<short>(), *uo = M->ju_off.as<short>();
    // L sweeps: y(m+1) = x - (L - I) y(m), y0 = x; the last one also writes z0 = y / d
    for (int m = 0; m < k; m++) {
        V *yn = ybuf[m & 1];
        const bool last = m == k - 1;
        if (m == 0) {
            if (last) sweep<V, double, double, V, true, true>(c, n, lp, li, lo, lv, dv, x, x, yn, zbuf[0], gate);
            else sweep<V, double, double, V, true, false>(c, n, lp, li, lo, lv, dv, x, x, yn, nullptr, gate);
        } else {
            const V *yo = ybuf[(m - 1) & 1];
            if (last) sweep<V, double, V, V, true, true>(c, n, lp, li, lo, lv, dv, x, yo, yn, zbuf[0], gate);
            else sweep<V, double, V, V, true, false>(c, n, lp, li, lo, lv, dv, x, yo, yn, nullptr, gate);
        }
    }
    // U sweeps: z(m+1) = (y - (U - D) z(m)) / d; the last writes the FP64 result
    const V *y = ybuf[(k - 1) & 1];
    for (int m = 0; m < k; m++) {
        const V *zo = zbuf[m & 1];
        if (m == k - 1) sweep<V, V, V, double, false, false>(c, n, up, ui, uo, uv, dv, y, zo, r, nullptr, gate);
        else sweep<V, V, V, V, false, false>(c, n, up, ui, uo, uv, dv, y, zo, zbuf[(m + 1) & 1], nullptr, gate);
    }
}

}  // namespace

void ilu_apply32x64_jacobi(hala_ilu M, const double *x, double *r, const int *gate, int k) {
    HALA_REQUIRE(k >= 1, HALA_ERR_ARG, "Jacobi ILU apply needs at least one sweep");
    if (!M->jac_values) {
        jac_values<float>(M, M->jl_val, M->ju_val, M->jd_val, M->jwork);
        M->jac_values = true;
    }
    jac_apply<float>(M, x, r, gate, k, M->jl_val.as<float>(), M->ju_val.as<float>(), M->jd_val.as<float>(),
                     M->jwork.as<float>());
}

void ilu_apply64_jacobi(hala_ilu M, const double *x, double *r, const int *gate, int k) {
    HALA_REQUIRE(k >= 1, HALA_ERR_ARG, "Jacobi ILU apply needs at least one sweep");
    if (!M->jac_values64) {
        jac_values<double>(M, M->jl_val64, M->ju_val64, M->jd_val64, M->jwork64);
        M->jac_values64 = true;
    }
    jac_apply<double>(M, x, r, gate, k, M->jl_val64.as<double>(), M->ju_val64.as<double>(),
                      M->jd_val64.as<double>(), M->jwork64.as<double>());
}

}  // namespace hala_amd
