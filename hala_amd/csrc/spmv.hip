// spmv.hip -- CSR handle, pattern validation and the SpMV kernels
// (replacing rocsparse_?csrmv of gpu/hala_rocm_sparse_general.hpp:115-128).
#include <algorithm>
#include <climits>

#include "spmv.hpp"

namespace hala_amd {

// ---------------------------------------------------------------------------
// row statistics and validation
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void row_stats_kernel(int nrows, const int *__restrict__ pntr, int *out) {
    __shared__ int smax[256], smin[256];
    int mx = 0, mn = INT_MAX;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nrows; i += gridDim.x * 256) {
        int l = pntr[i + 1] - pntr[i];
        mx = max(mx, l);
        mn = min(mn, l);
    }
    smax[threadIdx.x] = mx;
    smin[threadIdx.x] = mn;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (threadIdx.x < off) {
            smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + off]);
            smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax(out, smax[0]);
        atomicMin(out + 1, smin[0]);
    }
}

__global__ void init_stats(int *out) {
    out[0] = 0;
    out[1] = INT_MAX;
    out[2] = INT_MAX;
    out[3] = INT_MIN;
    out[4] = INT_MIN;
}

// out[3] = max (column - r0), out[4] = max (r0 - column) over the entries,
// r0 the first row of the entry's 256-row block (reduced over the wavefront,
// then an atomic only where it raises the stored value)
__global__ __launch_bounds__(256) void block_offset_kernel(int nrows, const int *__restrict__ pntr,
                                                           const int *__restrict__ indx, int *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int hi = INT_MIN, lo = INT_MIN;
    if (i < nrows) {
        const int r0 = i / SPMV_ROWS * SPMV_ROWS;
        for (int j = pntr[i]; j < pntr[i + 1]; j++) {
            hi = max(hi, indx[j] - r0);
            lo = max(lo, r0 - indx[j]);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        hi = max(hi, __shfl_xor(hi, o, 64));
        lo = max(lo, __shfl_xor(lo, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        if (hi > ld_agent(out + 3)) atomicMax(out + 3, hi);
        if (lo > ld_agent(out + 4)) atomicMax(out + 4, lo);
    }
}

__global__ void block_offset_fill(int nrows, const int *__restrict__ pntr, const int *__restrict__ indx,
                                  short *__restrict__ boff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const int r0 = i / SPMV_ROWS * SPMV_ROWS;
    for (int j = pntr[i]; j < pntr[i + 1]; j++) boff[j] = (short)(indx[j] - r0);
}

bool build_block_offsets(hala_ctx c, int nrows, int64_t nnz, const int *ptr, const int *idx, dbuf &out) {
    if (nrows <= 0 || nnz <= 0) return false;
    tmp_scope ts(c);
    dbuf &st = ts.get();
    st.reserve(5 * sizeof(int));
    init_stats<<<1, 1, 0, c->stream>>>(st.as<int>());
    block_offset_kernel<<<(nrows + 255) / 256, 256, 0, c->stream>>>(nrows, ptr, idx, st.as<int>());
    HALA_CHECK_LAUNCH();
    int h[5];
    HALA_HIP(hipMemcpyAsync(c->host_pinned, st.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HALA_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(h, c->host_pinned, sizeof(h));
    if (h[3] > SHRT_MAX || h[4] > -(int)SHRT_MIN) return false;
    out.reserve((size_t)nnz * sizeof(short));
    block_offset_fill<<<(nrows + 255) / 256, 256, 0, c->stream>>>(nrows, ptr, idx, out.as<short>());
    HALA_CHECK_LAUNCH();
    return true;
}

// out[2] = first bad row (INT_MAX if none): decreasing pointers, column out
// of range, unsorted/duplicate columns, or (require_diag) no diagonal entry.
__global__ void validate_kernel(int nrows, int ncols, int64_t nnz, const int *__restrict__ pntr,
                                const int *__restrict__ indx, int require_diag, int *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    int s = pntr[i], e = pntr[i + 1];
    bool bad = e < s || s < 0 || (int64_t)e > nnz || (i == 0 && s != 0) || (i == nrows - 1 && (int64_t)e != nnz);
    bool has_diag = false;
    int prev = -1;
    for (int j = s; j < e && !bad; j++) {
        int c = indx[j];
        if (c < 0 || c >= ncols || c <= prev) bad = true;
        if (c == i) has_diag = true;
        prev = c;
    }
    if (require_diag && !has_diag) bad = true;
    if (bad) atomicMin(out + 2, i);
}

int validate_pattern(hala_ctx c, int nrows, int ncols, int64_t nnz, const int *pntr, const int *indx,
                     int require_diag) {
    if (nrows == 0) return -1;
    int *st = reinterpret_cast<int *>(c->scal.as<double>() + S_DOT + 2);
    init_stats<<<1, 1, 0, c->stream>>>(st);
    validate_kernel<<<(nrows + 255) / 256, 256, 0, c->stream>>>(nrows, ncols, nnz, pntr, indx, require_diag, st);
    HALA_CHECK_LAUNCH();
    int r = fetch_i32(c, st + 2);
    return (r == INT_MAX) ? -1 : r;
}

// ---------------------------------------------------------------------------
// products
// ---------------------------------------------------------------------------
// y = alpha*sum + beta*y (beta == 0: y not read).  f32_inner: the FP32 sum is
// first normalised as 1*sum + 0 in FP32, then combined in the output type.
template <class TV, class TX, class TY, class TA, bool NT, class TI = int>
__global__ __launch_bounds__(256) void spmv_rows_kernel(int nrows, const int *__restrict__ pntr,
                                                        const TI *__restrict__ indx, const TV *__restrict__ vals,
                                                        const TX *__restrict__ x, TY alpha, TY beta,
                                                        TY *__restrict__ y, int f32_inner) {
    __shared__ TA prod[spmv_tile<TA>::value];
    TA acc = csr_block_rowsum<TV, TX, TA, false, NT, TI>(nrows, pntr, indx, vals, x, prod, blockIdx.x);
    const int row = blockIdx.x * SPMV_ROWS + threadIdx.x;
    if (row >= nrows) return;
    TY s;
    if (f32_inner) s = (TY)((TA)1 * acc + (TA)0);
    else s = (TY)acc;
    const TY out = (beta == TY(0)) ? alpha * s + TY(0) : alpha * s + beta * y[row];
    if (NT) __builtin_nontemporal_store(out, y + row);
    else y[row] = out;
}

// Value of v held by lane k (k wave-uniform), as a wave-uniform operand.
__device__ __forceinline__ double lane_value(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float lane_value(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// One wavefront per row for long rows.  The lanes form the products of 64
// consecutive entries at once (the next 64 are loaded while these are added);
// the sum then adds them one by one in CSR order, as a wave-uniform chain over
// the lanes' values: the reference's sequential row sum, bit for bit.
template <class TV, class TX, class TY, class TA>
__global__ __launch_bounds__(256) void spmv_wave_kernel(int nrows, const int *__restrict__ pntr,
                                                        const int *__restrict__ indx, const TV *__restrict__ vals,
                                                        const TX *__restrict__ x, TY alpha, TY beta,
                                                        TY *__restrict__ y, int f32_inner) {
    const int lane = threadIdx.x & 63;
    int row = (blockIdx.x * 256 + threadIdx.x) >> 6;
    const int nw = (gridDim.x * 256) >> 6;
    for (; row < nrows; row += nw) {
        const int s = pntr[row], e = pntr[row + 1];
        TA acc = TA(0);
        TA p = TA(0);
        if (s + lane < e) p = (TA)vals[s + lane] * (TA)x[indx[s + lane]];
        for (int base = s; base < e; base += 64) {
            const int jn = base + 64 + lane;
            TA pn = TA(0);
            if (jn < e) pn = (TA)vals[jn] * (TA)x[indx[jn]];
            const int cnt = min(64, e - base);
            for (int k = 0; k < cnt; k++) acc = acc + lane_value(p, k);
            p = pn;
        }
        if (lane == 0) {
            TY s = f32_inner ? (TY)((TA)1 * acc + (TA)0) : (TY)acc;
            y[row] = (beta == TY(0)) ? alpha * s + TY(0) : alpha * s + beta * y[row];
        }
    }
}

// Transposed product through the transposed pattern: y_k = beta*y_k (0 when
// beta == 0), then y_k += (alpha*x_i)*v for the entries of column k in
// increasing source row i -- the element order of the reference's scatter loop.
template <class T>
__global__ void spmv_t_kernel(int ncols, const int *__restrict__ tpntr, const int *__restrict__ trow,
                              const int *__restrict__ tmap, const T *__restrict__ vals, const T *__restrict__ x,
                              T alpha, T beta, T *__restrict__ y) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncols) return;
    T acc = (beta == T(0)) ? T(0) : beta * y[k];
    for (int e = tpntr[k]; e < tpntr[k + 1]; e++) acc = acc + (alpha * x[trow[e]]) * vals[tmap[e]];
    y[k] = acc;
}

template <class TV, class TX, class TY>
void spmv_axpby(hala_csr A, const TV *vals, const TX *x, TY alpha, TY beta, TY *y, bool f32_inner) {
    hala_ctx c = A->ctx;
    if (A->nrows == 0) return;
    using TA = TV;
    if (A->kernel == 1) {
        int g = grid_for((int64_t)A->nrows * 64, 256, 65536);
        spmv_wave_kernel<TV, TX, TY, TA><<<g, 256, 0, c->stream>>>(A->nrows, A->pntr, A->indx, vals, x, alpha, beta,
                                                                  y, f32_inner ? 1 : 0);
    } else {
        int g = (A->nrows + SPMV_ROWS - 1) / SPMV_ROWS;
        const int fi = f32_inner ? 1 : 0;
        const short *bo = A->boff.as<short>();
        if (stream_nt<TV>(A->nnz)) {
            if (bo)
                spmv_rows_kernel<TV, TX, TY, TA, true, short><<<g, SPMV_ROWS, 0, c->stream>>>(A->nrows, A->pntr, bo, vals,
                                                                                            x, alpha, beta, y, fi);
            else
                spmv_rows_kernel<TV, TX, TY, TA, true><<<g, SPMV_ROWS, 0, c->stream>>>(A->nrows, A->pntr, A->indx, vals,
                                                                                     x, alpha, beta, y, fi);
        } else {
            if (bo)
                spmv_rows_kernel<TV, TX, TY, TA, false, short><<<g, SPMV_ROWS, 0, c->stream>>>(A->nrows, A->pntr, bo,
                                                                                             vals, x, alpha, beta, y, fi);
            else
                spmv_rows_kernel<TV, TX, TY, TA, false><<<g, SPMV_ROWS, 0, c->stream>>>(A->nrows, A->pntr, A->indx,
                                                                                      vals, x, alpha, beta, y, fi);
        }
    }
    HALA_CHECK_LAUNCH();
}
template void spmv_axpby<double, double, double>(hala_csr, const double *, const double *, double, double, double *,
                                                 bool);
template void spmv_axpby<float, float, float>(hala_csr, const float *, const float *, float, float, float *, bool);
template void spmv_axpby<float, double, double>(hala_csr, const float *, const double *, double, double, double *,
                                                bool);

void ensure_transpose(hala_csr A) {
    if (A->tp.built) return;
    build_transpose(A->ctx, A->nrows, A->ncols, A->nnz, A->pntr, A->indx, A->tp);
}

template <class T>
void spmv_trans(hala_csr A, T alpha, const T *vals, const T *x, T beta, T *y) {
    ensure_transpose(A);
    hala_ctx c = A->ctx;
    if (A->ncols == 0) return;
    spmv_t_kernel<T><<<(A->ncols + 255) / 256, 256, 0, c->stream>>>(A->ncols, A->tp.tpntr.as<int>(),
                                                                    A->tp.trow.as<int>(), A->tp.tmap.as<int>(), vals,
                                                                    x, alpha, beta, y);
    HALA_CHECK_LAUNCH();
}
void spmm_cols(hala_csr A, char transa, char transb, int b_rows, int b_cols, size_t esize, const void *B, int ldb,
               void *C, int ldc, const std::function<void(const void *, void *)> &mv,
               const std::function<void(void *, int)> &conj) {
    auto isn = [](char t) { return t == 'N' || t == 'n'; };
    auto ok = [&](char t) { return isn(t) || t == 'T' || t == 't' || t == 'C' || t == 'c'; };
    HALA_REQUIRE(A && ok(transa) && ok(transb), HALA_ERR_ARG, "spmm: trans must be N, T or C");
    HALA_REQUIRE(b_rows >= 0 && b_cols >= 0, HALA_ERR_ARG, "spmm: negative size");
    const int M = isn(transa) ? A->nrows : A->ncols;  // rows of C
    const int K = isn(transa) ? A->ncols : A->nrows;  // rows of op(B)
    const int N = isn(transb) ? b_cols : b_rows;      // columns of C
    HALA_REQUIRE((isn(transb) ? b_rows : b_cols) == K, HALA_ERR_ARG, "spmm: op(B) must have op(A)'s column count of rows");
    HALA_REQUIRE(ldb >= std::max(b_rows, 1) && ldc >= std::max(M, 1), HALA_ERR_ARG, "spmm: ldb or ldc below the minimum");
    HALA_REQUIRE(N == 0 || (B && C), HALA_ERR_ARG, "spmm: null pointer");
    hala_ctx c = A->ctx;
    const char *b = static_cast<const char *>(B);
    char *cc = static_cast<char *>(C);
    tmp_scope ts(c);
    dbuf &t = ts.get();
    if (!isn(transb)) t.reserve((size_t)std::max(K, 1) * esize);
    for (int i = 0; i < N; i++) {
        char *y = cc + (size_t)i * (size_t)ldc * esize;
        if (isn(transb)) {
            mv(b + (size_t)i * (size_t)ldb * esize, y);
        } else {
            strided_copy(c, K, esize, b + (size_t)i * esize, ldb, t.p, 1);
            if ((transb == 'C' || transb == 'c') && conj) conj(t.p, K);
            mv(t.p, y);
        }
    }
}

template void spmv_trans<double>(hala_csr, double, const double *, const double *, double, double *);
template void spmv_trans<float>(hala_csr, float, const float *, const float *, float, float *);

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

int hala_csr_create(hala_ctx c, int nrows, int ncols, int64_t nnz, const int *pntr, const int *indx,
                    hala_csr *out) {
    return guard([&] {
        HALA_REQUIRE(c && out, HALA_ERR_ARG, "hala_csr_create: null argument");
        *out = nullptr;
        HALA_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0 && nnz < INT_MAX, HALA_ERR_ARG,
                     "hala_csr_create: bad sizes (int32 indexing)");
        HALA_REQUIRE(pntr && (indx || nnz == 0), HALA_ERR_ARG, "hala_csr_create: null pattern");
        set_device(c);
        auto *A = new hala_csr_s();
        A->ctx = c;
        A->nrows = nrows;
        A->ncols = ncols;
        A->nnz = nnz;
        A->pntr = pntr;
        A->indx = indx;
        if (nrows > 0) {
            // reject patterns that would send a kernel out of bounds
            int *vt = reinterpret_cast<int *>(c->scal.as<double>() + S_DOT + 2);
            init_stats<<<1, 1, 0, c->stream>>>(vt);
            validate_kernel<<<(nrows + 255) / 256, 256, 0, c->stream>>>(nrows, ncols, nnz, pntr, indx, 0, vt);
            hipError_t ev = hipGetLastError();
            int bad = INT_MAX;
            if (ev == hipSuccess) {
                try {
                    bad = fetch_i32(c, vt + 2);
                } catch (...) {
                    delete A;
                    throw;
                }
            }
            if (ev != hipSuccess || bad != INT_MAX) {
                delete A;
                if (ev != hipSuccess) throw status_error{HALA_ERR_HIP, std::string("csr validation: ") + hipGetErrorString(ev)};
                throw status_error{HALA_ERR_PATTERN, "hala_csr_create: row " + std::to_string(bad) +
                                                         " has decreasing pointers or unsorted/out-of-range columns"};
            }
            int *st = reinterpret_cast<int *>(c->scal.as<double>() + S_DOT + 2);
            init_stats<<<1, 1, 0, c->stream>>>(st);
            row_stats_kernel<<<grid_for(nrows, 256 * 8, 1024), 256, 0, c->stream>>>(nrows, pntr, st);
            block_offset_kernel<<<(nrows + 255) / 256, 256, 0, c->stream>>>(nrows, pntr, indx, st);
            int h[5];
            hipError_t e1 = hipGetLastError();
            if (e1 == hipSuccess) e1 = hipMemcpyAsync(c->host_pinned, st, 5 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
            if (e1 == hipSuccess) e1 = hipStreamSynchronize(c->stream);
            if (e1 != hipSuccess) {
                delete A;
                throw status_error{HALA_ERR_HIP, std::string("csr statistics: ") + hipGetErrorString(e1)};
            }
            std::memcpy(h, c->host_pinned, sizeof(h));
            A->max_row = h[0];
            A->min_row = h[1];
            A->mean_row = (double)nnz / nrows;
            // Short rows (every stencil here): LDS-staged blocks keep the
            // index/value streams fully coalesced and sum each row in CSR
            // order.  Long rows waste the per-row thread there: one
            // wavefront per row instead.
            A->kernel = (A->mean_row > 48.0 || A->max_row > 4 * spmv_tile<double>::value) ? 1 : 0;
            // banded patterns: columns as int16 offsets from the block's first row
            if (A->kernel == 0 && nnz > 0 && h[3] <= SHRT_MAX && h[4] <= -(int)SHRT_MIN) {
                try {
                    A->boff.reserve((size_t)nnz * sizeof(short));
                } catch (...) {
                    delete A;
                    throw;
                }
                block_offset_fill<<<(nrows + 255) / 256, 256, 0, c->stream>>>(nrows, pntr, indx, A->boff.as<short>());
                hipError_t e2 = hipGetLastError();
                if (e2 == hipSuccess) e2 = hipStreamSynchronize(c->stream);
                if (e2 != hipSuccess) {
                    delete A;
                    throw status_error{HALA_ERR_HIP, std::string("csr offsets: ") + hipGetErrorString(e2)};
                }
            }
        }
        *out = A;
    });
}

int hala_csr_destroy(hala_csr A) {
    return guard([&] {
        if (!A) return;
        (void)hipSetDevice(A->ctx->device);
        (void)hipStreamSynchronize(A->ctx->stream);
        delete A;
    });
}

int hala_csr_info(hala_csr A, hala_csr_info_t *info) {
    return guard([&] {
        HALA_REQUIRE(A && info, HALA_ERR_ARG, "hala_csr_info: null argument");
        info->nrows = A->nrows;
        info->ncols = A->ncols;
        info->nnz = A->nnz;
        info->max_row = A->max_row;
        info->min_row = A->min_row;
        info->mean_row = A->mean_row;
        info->kernel = A->kernel;
        info->index_bytes = A->boff.p ? 2 : 4;
    });
}

int hala_csr_set_kernel(hala_csr A, int kernel) {
    return guard([&] {
        HALA_REQUIRE(A && (kernel == 0 || kernel == 1), HALA_ERR_ARG, "hala_csr_set_kernel: bad argument");
        A->kernel = kernel;
    });
}

int hala_csr_validate(hala_csr A, int require_diag, int *bad_row) {
    return guard([&] {
        HALA_REQUIRE(A && bad_row, HALA_ERR_ARG, "hala_csr_validate: null argument");
        hala_ctx c = A->ctx;
        set_device(c);
        *bad_row = -1;
        if (A->nrows == 0) return;
        int *st = reinterpret_cast<int *>(c->scal.as<double>() + S_DOT + 2);
        init_stats<<<1, 1, 0, c->stream>>>(st);
        validate_kernel<<<(A->nrows + 255) / 256, 256, 0, c->stream>>>(A->nrows, A->ncols, A->nnz, A->pntr,
                                                                       A->indx, require_diag, st);
        HALA_CHECK_LAUNCH();
        int r = fetch_i32(c, st + 2);
        *bad_row = (r == INT_MAX) ? -1 : r;
    });
}

int hala_spmv_f64(hala_csr A, char trans, double alpha, const double *vals, const double *x, double beta,
                  double *y) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A, HALA_ERR_ARG, "spmv: null matrix");
        HALA_REQUIRE((vals || A->nnz == 0) && x && y, HALA_ERR_ARG, "spmv: null pointer");
        set_device(A->ctx);
        if (trans == 'N' || trans == 'n') spmv_axpby<double, double, double>(A, vals, x, alpha, beta, y, false);
        else if (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c')
            spmv_trans<double>(A, alpha, vals, x, beta, y);
        else throw status_error{HALA_ERR_ARG, "spmv: trans must be N, T or C"};
    });
}

int hala_spmv_f32(hala_csr A, char trans, float alpha, const float *vals, const float *x, float beta, float *y) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A, HALA_ERR_ARG, "spmv: null matrix");
        HALA_REQUIRE((vals || A->nnz == 0) && x && y, HALA_ERR_ARG, "spmv: null pointer");
        set_device(A->ctx);
        if (trans == 'N' || trans == 'n') spmv_axpby<float, float, float>(A, vals, x, alpha, beta, y, false);
        else if (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c')
            spmv_trans<float>(A, alpha, vals, x, beta, y);
        else throw status_error{HALA_ERR_ARG, "spmv: trans must be N, T or C"};
    });
}

int hala_spmm_f64(hala_csr A, char transa, char transb, int b_rows, int b_cols, double alpha, const double *vals,
                  const double *B, int ldb, double beta, double *C, int ldc) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A, HALA_ERR_ARG, "spmm: null matrix");
        set_device(A->ctx);
        spmm_cols(A, transa, transb, b_rows, b_cols, sizeof(double), B, ldb, C, ldc,
                  [&](const void *x, void *y) {
                      const int rc = hala_spmv_f64(A, transa, alpha, vals, (const double *)x, beta, (double *)y);
                      if (rc != HALA_OK) throw status_error{rc, hala_last_error()};
                  },
                  nullptr);
    });
}

int hala_spmm_f32(hala_csr A, char transa, char transb, int b_rows, int b_cols, float alpha, const float *vals,
                  const float *B, int ldb, float beta, float *C, int ldc) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A, HALA_ERR_ARG, "spmm: null matrix");
        set_device(A->ctx);
        spmm_cols(A, transa, transb, b_rows, b_cols, sizeof(float), B, ldb, C, ldc,
                  [&](const void *x, void *y) {
                      const int rc = hala_spmv_f32(A, transa, alpha, vals, (const float *)x, beta, (float *)y);
                      if (rc != HALA_OK) throw status_error{rc, hala_last_error()};
                  },
                  nullptr);
    });
}

int hala_spmv_f32x64(hala_csr A, double alpha, const float *vals, const double *x, double beta, double *y) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A, HALA_ERR_ARG, "spmv: null matrix");
        HALA_REQUIRE((vals || A->nnz == 0) && x && y, HALA_ERR_ARG, "spmv: null pointer");
        set_device(A->ctx);
        spmv_axpby<float, double, double>(A, vals, x, alpha, beta, y, true);
    });
}

}  // extern "C"
