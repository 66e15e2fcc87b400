// cplx.hip -- complex scalars (std::complex<double> / <float> layout: real,
// imaginary interleaved) on the solve path: BLAS-1, the CSR product ('N', 'T',
// 'C'), ILU(0) factorisation and its two triangular sweeps.  The reference
// instantiates every routine for the four scalar types (tests/solvers_tests.cpp:
// 14-40); here the complex ones are plain level-by-level kernels (the real FP64
// path carries the tuned sweeps).
//
// Arithmetic, never fused (-ffp-contract=off):
//   a * b = (ar br - ai bi, ar bi + ai br)
//   a / b = ((ar br + ai bi) / d, (ai br - ar bi) / d),  d = br br + bi bi
//   conj-dot terms (dotc): (xr yr + xi yi, xr yi - xi yr); dotu: (xr yr - xi yi, xr yi + xi yr)
// Reductions sum those per-element terms in the canonical order of the real
// reductions (256-element blocks, halving tree, sequential pass over the
// block partials, final tree), real and imaginary parts separately.
#include <algorithm>
#include <climits>

#include "ilu.hpp"
#include "spmv.hpp"

namespace hala_amd {

template <class R>
struct cx {
    R re, im;
};
template <class R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) {
    return {a.re + b.re, a.im + b.im};
}
template <class R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) {
    return {a.re - b.re, a.im - b.im};
}
template <class R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__device__ __forceinline__ cx<R> cdiv(cx<R> a, cx<R> b) {
    const R d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
template <class R>
__device__ __forceinline__ cx<R> cconj(cx<R> a) {
    return {a.re, -a.im};
}
template <class R>
__device__ __forceinline__ cx<R> ld(const R *p, int64_t i) {
    return {p[2 * i], p[2 * i + 1]};
}
template <class R>
__device__ __forceinline__ void st(R *p, int64_t i, cx<R> v) {
    p[2 * i] = v.re;
    p[2 * i + 1] = v.im;
}

namespace {

enum zterm : int { Z_DOTC_RE = 0, Z_DOTC_IM, Z_DOTU_RE, Z_DOTU_IM, Z_NRM2 };

template <class R>
__device__ __forceinline__ R term(int mode, const R *x, const R *y, int64_t i) {
    const R xr = x[2 * i], xi = x[2 * i + 1];
    if (mode == Z_NRM2) return xr * xr + xi * xi;
    const R yr = y[2 * i], yi = y[2 * i + 1];
    switch (mode) {
    case Z_DOTC_RE: return xr * yr + xi * yi;
    case Z_DOTC_IM: return xr * yi - xi * yr;
    case Z_DOTU_RE: return xr * yr - xi * yi;
    default: return xr * yi + xi * yr;
    }
}

// one real canonical reduction of the per-element terms (the structure of cdot_kernel)
template <class R, int K>
__global__ __launch_bounds__(256) void zdot_kernel(int64_t n, const R *__restrict__ x, const R *__restrict__ y,
                                                   int64_t nb, R *part, R *out, int mode, int *err) {
    __shared__ R sh[CB * K];
    R v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        int64_t i = ((int64_t)blockIdx.x * K + k) * CB + threadIdx.x;
        v[k] = i < n ? term<R>(mode, x, y, i) : R(0);
    }
    tree256<R, K>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            int64_t b = (int64_t)blockIdx.x * K + k;
            if (b < nb) st_agent(part + b, v[k]);
        }
    }
    if (is_collector()) {
        R r = final_pass<R>(part, nb, sh, err);
        if (threadIdx.x == 0) out[0] = mode == Z_NRM2 ? (R)sqrt(r) : r;
    }
}

template <class R>
void zreduce(hala_ctx c, int64_t n, const R *x, const R *y, R *out_dev, int mode) {
    constexpr int K = 4;
    int64_t nb = std::max<int64_t>((n + CB - 1) / CB, 1);
    R *part;
    if constexpr (sizeof(R) == 8) part = reinterpret_cast<R *>(partials64(c, nb));
    else part = reinterpret_cast<R *>(partials32(c, nb));
    int64_t g = (nb + K - 1) / K;
    HALA_REQUIRE(g < (1ll << 31), HALA_ERR_ARG, "vector too long");
    zdot_kernel<R, K><<<(unsigned)g, CB, 0, c->stream>>>(n, x, y, nb, part, out_dev, mode, wait_error_flag(c));
    HALA_CHECK_LAUNCH();
}

// res[0..nres) on the host: the reductions of modes m0 (and m1)
template <class R>
void zdot_host(hala_ctx c, int64_t n, const R *x, const R *y, R *res, int m0, int m1) {
    HALA_REQUIRE(c && n >= 0 && res, HALA_ERR_ARG, "complex reduction: bad argument");
    HALA_REQUIRE(n == 0 || (x && (y || m0 == Z_NRM2)), HALA_ERR_ARG, "complex reduction: null vector");
    set_device(c);
    const int nres = m1 < 0 ? 1 : 2;
    if (n == 0) {
        for (int k = 0; k < nres; k++) res[k] = R(0);
        return;
    }
    R *out = reinterpret_cast<R *>(c->scal.as<double>() + S_DOT);
    zreduce<R>(c, n, x, y, out, m0);
    if (m1 >= 0) zreduce<R>(c, n, x, y, out + 1, m1);
    HALA_HIP(hipMemcpyAsync(c->host_pinned, out, nres * sizeof(R), hipMemcpyDeviceToHost, c->stream));
    HALA_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < nres; k++) res[k] = reinterpret_cast<R *>(c->host_pinned)[k];
}

template <class R>
__global__ void zaxpy_kernel(int64_t n, cx<R> a, const R *__restrict__ x, R *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) st(y, i, cadd(ld(y, i), cmul(a, ld(x, i))));
}
template <class R>
__global__ void zscal_kernel(int64_t n, cx<R> a, R *__restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) st(x, i, cmul(a, ld(x, i)));
}

// y_i = alpha * (sum_j A_ij x_j, j in CSR order) + beta * y_i  (beta == 0: y not read)
template <class R>
__global__ void zspmv_n_kernel(int nrows, const int *__restrict__ pntr, const int *__restrict__ indx,
                               const R *__restrict__ vals, const R *__restrict__ x, cx<R> alpha, cx<R> beta,
                               R *__restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    cx<R> acc = {R(0), R(0)};
    for (int j = pntr[i]; j < pntr[i + 1]; j++) acc = cadd(acc, cmul(ld(vals, j), ld(x, indx[j])));
    cx<R> out = cmul(alpha, acc);
    if (beta.re != R(0) || beta.im != R(0)) out = cadd(out, cmul(beta, ld(y, i)));
    st(y, i, out);
}

// y_k = beta y_k (0 when beta == 0), then y_k += (alpha x_i) * op(A_ik) for the entries
// of column k in increasing source row i (the reference's scatter order); op = conj for 'C'
template <class R>
__global__ void zspmv_t_kernel(int ncols, const int *__restrict__ tpntr, const int *__restrict__ trow,
                               const int *__restrict__ tmap, const R *__restrict__ vals, const R *__restrict__ x,
                               cx<R> alpha, cx<R> beta, int conj, R *__restrict__ y) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncols) return;
    cx<R> acc = (beta.re != R(0) || beta.im != R(0)) ? cmul(beta, ld(y, k)) : cx<R>{R(0), R(0)};
    for (int e = tpntr[k]; e < tpntr[k + 1]; e++) {
        cx<R> v = ld(vals, tmap[e]);
        if (conj) v = cconj(v);
        acc = cadd(acc, cmul(cmul(alpha, ld(x, trow[e])), v));
    }
    st(y, k, acc);
}

// IKJ ILU(0) of the rows of one level (the operation order of the real factor in ilu.hip, complex arithmetic)
template <class R>
__global__ void zilu_level_kernel(const int *__restrict__ rows, int count, const int *__restrict__ pntr,
                                  const int *__restrict__ indx, const int *__restrict__ diag, R *f) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int j = rows[t], je = pntr[j + 1], dj = diag[j];
    for (int jc = pntr[j]; jc < dj; jc++) {
        const int i = indx[jc];
        const cx<R> l = cdiv(ld(f, jc), ld(f, diag[i]));
        st(f, jc, l);
        int ik = diag[i] + 1, jk = jc + 1;
        const int ie = pntr[i + 1];
        while (ik < ie && jk < je) {
            const int ci = indx[ik], cj = indx[jk];
            if (ci == cj) {
                st(f, jk, csub(ld(f, jk), cmul(l, ld(f, ik))));
                ik++;
                jk++;
            } else if (ci < cj) {
                ik++;
            } else {
                jk++;
            }
        }
    }
}

// one level of the forward sweep (unit L): y_i = x_i - l_ij y_j ..., j < i in row order
template <class R>
__global__ void zlower_level(const int *__restrict__ rows, int count, const int *__restrict__ pntr,
                             const int *__restrict__ indx, const int *__restrict__ diag, const R *__restrict__ f,
                             const R *__restrict__ x, R *__restrict__ y) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = rows[t];
    cx<R> acc = ld(x, i);
    for (int j = pntr[i]; j < diag[i]; j++) acc = csub(acc, cmul(ld(f, j), ld(y, indx[j])));
    st(y, i, acc);
}

// one level of the backward sweep: z_i = (y_i - u_ij z_j ...) / u_ii, j > i in row order
template <class R>
__global__ void zupper_level(const int *__restrict__ rows, int count, const int *__restrict__ pntr,
                             const int *__restrict__ indx, const int *__restrict__ diag, const R *__restrict__ f,
                             const R *__restrict__ y, R *__restrict__ z) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = rows[t], d = diag[i];
    cx<R> acc = ld(y, i);
    for (int j = d + 1; j < pntr[i + 1]; j++) acc = csub(acc, cmul(ld(f, j), ld(z, indx[j])));
    st(z, i, cdiv(acc, ld(f, d)));
}

// One level of a standalone triangular solve through a plan's dependency
// lists: x_r = (alpha b_r - sum_k v_k x_dep(k)) [/ d_r], the terms in the
// plan's order (row order for 'N', the reference's source-row order for 'T'
// and 'C').  CONJ ('C'): conjugated values and diagonal.
template <class R, bool CONJ>
__global__ void ztri_level(const int *__restrict__ rows, int count, const int *__restrict__ dptr,
                           const int *__restrict__ ddep, const int *__restrict__ dsrc, const int *__restrict__ dgs,
                           const R *__restrict__ vals, cx<R> alpha, const R *__restrict__ b, R *__restrict__ x) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int r = rows[t];
    cx<R> s = {R(0), R(0)};
    for (int k = dptr[r]; k < dptr[r + 1]; k++) {
        cx<R> v = ld(vals, dsrc[k]);
        if (CONJ) v = cconj(v);
        s = cadd(s, cmul(v, ld(x, ddep[k])));
    }
    cx<R> y = csub(cmul(alpha, ld(b, r)), s);
    const int dp = dgs[r];
    if (dp >= 0) {
        cx<R> d = ld(vals, dp);
        if (CONJ) d = cconj(d);
        y = cdiv(y, d);
    }
    st(x, r, y);
}

template <class R>
__global__ void zconj_kernel(int n, R *__restrict__ x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[2 * (int64_t)i + 1] = -x[2 * (int64_t)i + 1];
}

template <class R>
__global__ void zpivot_check(int n, const int *__restrict__ diag, const R *__restrict__ f, int *bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R ur = f[2 * (int64_t)diag[i]], ui = f[2 * (int64_t)diag[i] + 1];
    if ((ur == R(0) && ui == R(0)) || !isfinite(ur) || !isfinite(ui)) atomicMin(bad, i);
}

__global__ void set_i32_k(int *p, int v) { *p = v; }

template <class R>
void launch_1d_z(hala_ctx c, int64_t n, void (*k)(int64_t, cx<R>, const R *, R *), cx<R> a, const R *x, R *y) {
    if (n <= 0) return;
    k<<<grid_for(n, 256 * 4, 16384), 256, 0, c->stream>>>(n, a, x, y);
    HALA_CHECK_LAUNCH();
}

template <class R>
dbuf &zfactor(hala_ilu M) {
    return sizeof(R) == 8 ? M->zfactor : M->cfactor;
}

template <class R>
void zilu_factor(hala_ilu M, const R *vals) {
    hala_ctx c = M->ctx;
    set_device(c);
    dbuf &F = zfactor<R>(M);
    if (!F.p) F.reserve((size_t)M->nnz * 2 * sizeof(R));
    R *f = F.as<R>();
    HALA_HIP(hipMemcpyAsync(f, vals, (size_t)M->nnz * 2 * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
    const std::vector<int> &lp = M->L.h_level_ptr;
    const int *perm = M->L.perm.as<int>();
    for (int l = 1; l < M->L.nlev; l++) {
        const int cnt = lp[l + 1] - lp[l];
        if (cnt <= 0) continue;
        zilu_level_kernel<R><<<(cnt + 255) / 256, 256, 0, c->stream>>>(perm + lp[l], cnt, M->pntr, M->indx,
                                                                       M->diag.as<int>(), f);
        HALA_CHECK_LAUNCH();
    }
    dbuf bad;
    bad.reserve(sizeof(int));
    set_i32_k<<<1, 1, 0, c->stream>>>(bad.as<int>(), INT_MAX);
    zpivot_check<R><<<(M->n + 255) / 256, 256, 0, c->stream>>>(M->n, M->diag.as<int>(), f, bad.as<int>());
    HALA_CHECK_LAUNCH();
    const int b = fetch_i32(c, bad.as<int>());
    if (sizeof(R) == 8) M->zfactored = false;
    else M->cfactored = false;
    HALA_REQUIRE(b == INT_MAX, HALA_ERR_PIVOT, "complex ILU: zero or non-finite pivot in row " + std::to_string(b));
    if (sizeof(R) == 8) M->zfactored = true;
    else M->cfactored = true;
}

template <class R>
void zilu_apply(hala_ilu M, const R *x, R *r) {
    hala_ctx c = M->ctx;
    set_device(c);
    HALA_REQUIRE(sizeof(R) == 8 ? M->zfactored : M->cfactored, HALA_ERR_ARG, "complex ILU apply before factorisation");
    const R *f = zfactor<R>(M).template as<R>();
    dbuf tmp;
    tmp.reserve((size_t)M->n * 2 * sizeof(R));
    R *y = tmp.as<R>();
    const int *d = M->diag.as<int>();
    {
        const std::vector<int> &lp = M->L.h_level_ptr;
        const int *perm = M->L.perm.as<int>();
        for (int l = 0; l < M->L.nlev; l++) {
            const int cnt = lp[l + 1] - lp[l];
            if (cnt <= 0) continue;
            zlower_level<R><<<(cnt + 255) / 256, 256, 0, c->stream>>>(perm + lp[l], cnt, M->pntr, M->indx, d, f, x,
                                                                      y);
        }
    }
    {
        const std::vector<int> &lp = M->U.h_level_ptr;
        const int *perm = M->U.perm.as<int>();
        for (int l = 0; l < M->U.nlev; l++) {
            const int cnt = lp[l + 1] - lp[l];
            if (cnt <= 0) continue;
            zupper_level<R><<<(cnt + 255) / 256, 256, 0, c->stream>>>(perm + lp[l], cnt, M->pntr, M->indx, d, f, y,
                                                                      r);
        }
    }
    HALA_CHECK_LAUNCH();
    HALA_HIP(hipStreamSynchronize(c->stream));
}

template <class R>
void ztrsv(hala_tri T, char trans, cx<R> alpha, const R *vals, const R *b, R *x) {
    HALA_REQUIRE(T && vals && b && x, HALA_ERR_ARG, "complex trsv: null argument");
    const bool tr = !(trans == 'N' || trans == 'n');
    const bool cj = trans == 'C' || trans == 'c';
    HALA_REQUIRE(!tr || cj || trans == 'T' || trans == 't', HALA_ERR_ARG, "complex trsv: trans must be N, T or C");
    hala_ctx c = T->ctx;
    set_device(c);
    tri_plan &p = tr ? T->T : T->N;
    const std::vector<int> &lp = p.h_level_ptr;
    const int *perm = p.perm.as<int>();
    for (int l = 0; l < p.nlev; l++) {
        const int cnt = lp[l + 1] - lp[l];
        if (cnt <= 0) continue;
        const int g = (cnt + 255) / 256;
        if (cj)
            ztri_level<R, true><<<g, 256, 0, c->stream>>>(perm + lp[l], cnt, p.dptr.as<int>(), p.ddep.as<int>(),
                                                          p.dsrc.as<int>(), p.diag_src.as<int>(), vals, alpha, b, x);
        else
            ztri_level<R, false><<<g, 256, 0, c->stream>>>(perm + lp[l], cnt, p.dptr.as<int>(), p.ddep.as<int>(),
                                                           p.dsrc.as<int>(), p.diag_src.as<int>(), vals, alpha, b, x);
    }
    HALA_CHECK_LAUNCH();
}

template <class R>
void zspmv(hala_csr A, char trans, cx<R> alpha, const R *vals, const R *x, cx<R> beta, R *y) {
    HALA_REQUIRE(A && (A->nnz == 0 || vals) && x && y, HALA_ERR_ARG, "complex spmv: null argument");
    HALA_REQUIRE(trans == 'N' || trans == 'n' || trans == 'T' || trans == 't' || trans == 'C' || trans == 'c',
                 HALA_ERR_ARG, "complex spmv: trans must be N, T or C");
    hala_ctx c = A->ctx;
    set_device(c);
    if (trans == 'N' || trans == 'n') {
        if (A->nrows == 0) return;
        zspmv_n_kernel<R><<<(A->nrows + 255) / 256, 256, 0, c->stream>>>(A->nrows, A->pntr, A->indx, vals, x, alpha,
                                                                         beta, y);
    } else {
        ensure_transpose(A);
        if (A->ncols == 0) return;
        zspmv_t_kernel<R><<<(A->ncols + 255) / 256, 256, 0, c->stream>>>(
            A->ncols, A->tp.tpntr.as<int>(), A->tp.trow.as<int>(), A->tp.tmap.as<int>(), vals, x, alpha, beta,
            (trans == 'C' || trans == 'c') ? 1 : 0, y);
    }
    HALA_CHECK_LAUNCH();
}

}  // namespace
}  // namespace hala_amd

using namespace hala_amd;

#define HALA_Z_ENTRIES(SUF, R)                                                                                     \
    int hala_dotc_##SUF(hala_ctx c, int64_t n, const R *x, const R *y, R *res) {                                  \
        return guard([&] { zdot_host<R>(c, n, x, y, res, Z_DOTC_RE, Z_DOTC_IM); });                              \
    }                                                                                                            \
    int hala_dotu_##SUF(hala_ctx c, int64_t n, const R *x, const R *y, R *res) {                                  \
        return guard([&] { zdot_host<R>(c, n, x, y, res, Z_DOTU_RE, Z_DOTU_IM); });                              \
    }                                                                                                            \
    int hala_nrm2_##SUF(hala_ctx c, int64_t n, const R *x, R *res) {                                             \
        return guard([&] { zdot_host<R>(c, n, x, nullptr, res, Z_NRM2, -1); });                                  \
    }                                                                                                            \
    int hala_axpy_##SUF(hala_ctx c, int64_t n, R are, R aim, const R *x, R *y) {                                  \
        return guard([&] {                                                                                       \
            HALA_REQUIRE(c && n >= 0 && (n == 0 || (x && y)), HALA_ERR_ARG, "complex axpy: bad argument");        \
            set_device(c);                                                                                       \
            launch_1d_z<R>(c, n, zaxpy_kernel<R>, cx<R>{are, aim}, x, y);                                         \
        });                                                                                                      \
    }                                                                                                            \
    int hala_scal_##SUF(hala_ctx c, int64_t n, R are, R aim, R *x) {                                              \
        return guard([&] {                                                                                       \
            HALA_REQUIRE(c && n >= 0 && (n == 0 || x), HALA_ERR_ARG, "complex scal: bad argument");               \
            set_device(c);                                                                                       \
            if (n > 0) {                                                                                         \
                zscal_kernel<R><<<grid_for(n, 256 * 4, 16384), 256, 0, c->stream>>>(n, cx<R>{are, aim}, x);       \
                HALA_CHECK_LAUNCH();                                                                             \
            }                                                                                                    \
        });                                                                                                      \
    }                                                                                                            \
    int hala_spmv_##SUF(hala_csr A, char trans, R are, R aim, const R *vals, const R *x, R bre, R bim, R *y) {   \
        return guard([&] { zspmv<R>(A, trans, cx<R>{are, aim}, vals, x, cx<R>{bre, bim}, y); });                 \
    }                                                                                                            \
    int hala_ilu_factor_##SUF(hala_ilu M, const R *vals) {                                                        \
        return guard([&] {                                                                                       \
            HALA_REQUIRE(M && vals, HALA_ERR_ARG, "complex ILU factor: null argument");                           \
            zilu_factor<R>(M, vals);                                                                             \
        });                                                                                                      \
    }                                                                                                            \
    int hala_ilu_apply_##SUF(hala_ilu M, const R *x, R *r) {                                                      \
        return guard([&] {                                                                                       \
            HALA_REQUIRE(M && x && r, HALA_ERR_ARG, "complex ILU apply: null argument");                          \
            zilu_apply<R>(M, x, r);                                                                              \
        });                                                                                                      \
    }                                                                                                            \
    int hala_spmm_##SUF(hala_csr A, char transa, char transb, int b_rows, int b_cols, R are, R aim, const R *vals, \
                        const R *B, int ldb, R bre, R bim, R *C, int ldc) {                                        \
        return guard([&] {                                                                                       \
            HALA_REQUIRE(A, HALA_ERR_ARG, "complex spmm: null matrix");                                           \
            set_device(A->ctx);                                                                                  \
            hala_ctx cx_ = A->ctx;                                                                               \
            spmm_cols(                                                                                           \
                A, transa, transb, b_rows, b_cols, 2 * sizeof(R), B, ldb, C, ldc,                                \
                [&](const void *x, void *y) {                                                                    \
                    zspmv<R>(A, transa, cx<R>{are, aim}, vals, (const R *)x, cx<R>{bre, bim}, (R *)y);           \
                },                                                                                               \
                [&](void *t, int n) {                                                                            \
                    if (n > 0) zconj_kernel<R><<<(n + 255) / 256, 256, 0, cx_->stream>>>(n, (R *)t);             \
                    HALA_CHECK_LAUNCH();                                                                         \
                });                                                                                              \
        });                                                                                                      \
    }                                                                                                            \
    int hala_trsv_##SUF(hala_tri T, char trans, R are, R aim, const R *vals, const R *b, R *x) {                 \
        return guard([&] { ztrsv<R>(T, trans, cx<R>{are, aim}, vals, b, x); });                                  \
    }                                                                                                            \
    int hala_trsm_##SUF(hala_tri T, char transa, char transb, int nrhs, R are, R aim, const R *vals, R *B,       \
                        int ldb) {                                                                               \
        return guard([&] {                                                                                       \
            HALA_REQUIRE(T && vals, HALA_ERR_ARG, "complex trsm: null argument");                                 \
            set_device(T->ctx);                                                                                  \
            tri_trsm_cols(T, transb, nrhs, 2 * sizeof(R), B, ldb, [&](const void *in, void *out) {                \
                ztrsv<R>(T, transa, cx<R>{are, aim}, vals, (const R *)in, (R *)out);                              \
            });                                                                                                  \
        });                                                                                                      \
    }

extern "C" {
HALA_Z_ENTRIES(z, double)
HALA_Z_ENTRIES(c, float)
}
