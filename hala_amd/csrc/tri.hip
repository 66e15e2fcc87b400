// tri.hip -- standalone sparse triangular matrix (gpu_triangular_matrix of
// gpu/hala_rocm_sparse_triangular.hpp:29-194) with the semantics of
// sparse_trsv_array (sparse/hala_sparse_utils.hpp:284-335): the CSR holds one
// triangle, the diagonal being the last ('L') or first ('U') entry of a row.
#include <climits>

#include "ilu.hpp"

namespace hala_amd {

// row's diagonal is the last (lower) / first (upper) entry and equals the row index
__global__ void tri_check(int n, int lower, const int *__restrict__ pntr, const int *__restrict__ indx,
                          int *__restrict__ diagpos, int *bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = pntr[i], e = pntr[i + 1];
    bool ok = e > s;
    int d = -1;
    if (ok) {
        d = lower ? e - 1 : s;
        ok = indx[d] == i;
        if (lower) ok = ok && (s == d || indx[s] < i);
        else ok = ok && (e - 1 == d || indx[e - 1] > i);
    }
    diagpos[i] = d < 0 ? s : d;
    if (!ok) atomicMin(bad, i);
}

__global__ void set_i32_tri(int *p, int v) { *p = v; }

__global__ void f32_to_f64(int64_t n, const float *__restrict__ a, double *__restrict__ b) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

// n elements of E from src (stride ss elements) to dst (stride ds)
template <class E>
__global__ void strided_copy_k(int n, const E *__restrict__ src, int64_t ss, E *__restrict__ dst, int64_t ds) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[(int64_t)i * ds] = src[(int64_t)i * ss];
}

void strided_copy(hala_ctx c, int n, size_t esize, const void *src, int64_t ss, void *dst, int64_t ds) {
    const int g = (n + 255) / 256;
    if (esize == 4)
        strided_copy_k<int><<<g, 256, 0, c->stream>>>(n, (const int *)src, ss, (int *)dst, ds);
    else if (esize == 8)
        strided_copy_k<int2><<<g, 256, 0, c->stream>>>(n, (const int2 *)src, ss, (int2 *)dst, ds);
    else
        strided_copy_k<int4><<<g, 256, 0, c->stream>>>(n, (const int4 *)src, ss, (int4 *)dst, ds);
    HALA_CHECK_LAUNCH();
}

static bool is_n(char t) { return t == 'N' || t == 'n'; }

void tri_trsm_cols(hala_tri T, char transb, int nrhs, size_t esize, void *B, int ldb,
                   const std::function<void(const void *, void *)> &col) {
    HALA_REQUIRE(T && (B || nrhs == 0) && nrhs >= 0, HALA_ERR_ARG, "hala_trsm: bad argument");
    HALA_REQUIRE(is_n(transb) || transb == 'T' || transb == 't' || transb == 'C' || transb == 'c', HALA_ERR_ARG,
                 "hala_trsm: transb must be N, T or C");
    const int n = T->n;
    const bool bn = is_n(transb);
    if (ldb < 0) ldb = bn ? n : nrhs;
    HALA_REQUIRE(ldb >= (bn ? n : nrhs) && ldb >= 1, HALA_ERR_ARG, "hala_trsm: ldb below the minimum");
    hala_ctx c = T->ctx;
    tmp_scope ts(c);
    dbuf &t1 = ts.get();
    dbuf &t2 = ts.get();
    t1.reserve((size_t)n * esize);
    t2.reserve((size_t)n * esize);
    char *b = static_cast<char *>(B);
    for (int i = 0; i < nrhs; i++) {
        if (bn) {  // column i is contiguous: a copy of it is the right-hand side, the solution lands in place
            char *bc = b + (size_t)i * (size_t)ldb * esize;
            HALA_HIP(hipMemcpyAsync(t1.p, bc, (size_t)n * esize, hipMemcpyDeviceToDevice, c->stream));
            col(t1.p, bc);
        } else {  // column i is row i of the stored B: gather, solve, scatter
            strided_copy(c, n, esize, b + (size_t)i * esize, ldb, t1.p, 1);
            col(t1.p, t2.p);
            strided_copy(c, n, esize, t2.p, 1, b + (size_t)i * esize, ldb);
        }
    }
}

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

int hala_tri_create(hala_ctx c, char uplo, char diag, int n, int64_t nnz, const int *pntr, const int *indx,
                    hala_tri *out) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(c && out, HALA_ERR_ARG, "hala_tri_create: null argument");
        *out = nullptr;
        bool lower = (uplo == 'L' || uplo == 'l');
        HALA_REQUIRE(lower || uplo == 'U' || uplo == 'u', HALA_ERR_ARG, "hala_tri_create: uplo must be L or U");
        bool unit = (diag == 'U' || diag == 'u');
        HALA_REQUIRE(unit || diag == 'N' || diag == 'n', HALA_ERR_ARG, "hala_tri_create: diag must be U or N");
        HALA_REQUIRE(n > 0 && nnz > 0 && nnz < INT_MAX && pntr && indx, HALA_ERR_ARG, "hala_tri_create: bad sizes");
        set_device(c);
        int bad = validate_pattern(c, n, n, nnz, pntr, indx, 1);
        HALA_REQUIRE(bad < 0, HALA_ERR_PATTERN, "hala_tri_create: invalid pattern at row " + std::to_string(bad));
        auto *T = new hala_tri_s();
        try {
            T->ctx = c;
            T->n = n;
            T->nnz = nnz;
            T->pntr = pntr;
            T->indx = indx;
            T->uplo = lower ? 'L' : 'U';
            T->diag = unit ? 'U' : 'N';
            dbuf dpos;
            dpos.reserve((size_t)n * sizeof(int));
            int *flag = reinterpret_cast<int *>(c->scal.as<double>() + S_DOT + 7);
            set_i32_tri<<<1, 1, 0, c->stream>>>(flag, INT_MAX);
            tri_check<<<(n + 255) / 256, 256, 0, c->stream>>>(n, lower ? 1 : 0, pntr, indx, dpos.as<int>(), flag);
            HALA_CHECK_LAUNCH();
            int b = fetch_i32(c, flag);
            HALA_REQUIRE(b == INT_MAX, HALA_ERR_PATTERN,
                         "hala_tri_create: row " + std::to_string(b) + " is not " +
                             (lower ? "lower triangular with the diagonal last" : "upper triangular with the diagonal first"));
            build_deps(c, T->N, n, lower ? DEP_ROW_BUT_LAST : DEP_ROW_BUT_FIRST, pntr, indx, dpos.as<int>(), nullptr,
                       unit, !lower);
            build_schedule(c, T->N, false);
            // transposed solve: L^T is upper (backward, source rows descending), U^T lower (forward, ascending)
            build_transpose(c, n, n, nnz, pntr, indx, T->tp);
            build_deps(c, T->T, n, lower ? DEP_T_UPPER : DEP_T_LOWER, pntr, indx, dpos.as<int>(), &T->tp, unit, lower);
            build_schedule(c, T->T, false);
            T->t_built = true;
        } catch (...) {
            delete T;
            throw;
        }
        *out = T;
    });
}

int hala_tri_destroy(hala_tri T) {
    HALA_RANGE();
    return guard([&] {
        if (!T) return;
        (void)hipSetDevice(T->ctx->device);
        (void)hipStreamSynchronize(T->ctx->stream);
        delete T;
    });
}

// x = op(T)^-1 (alpha b): sum of the off-diagonal terms first, then
// (alpha*b_i - s) [/ diag]; the transposed sums add the terms in the source-row
// order of the reference's column sweep.
int hala_trsv_f64(hala_tri T, char trans, double alpha, const double *vals, const double *b, double *x) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(T && vals && b && x, HALA_ERR_ARG, "hala_trsv_f64: null argument");
        hala_ctx c = T->ctx;
        set_device(c);
        bool tr = !(trans == 'N' || trans == 'n');
        HALA_REQUIRE(!tr || trans == 'T' || trans == 't' || trans == 'C' || trans == 'c', HALA_ERR_ARG,
                     "hala_trsv_f64: trans must be N, T or C");
        tri_plan &p = tr ? T->T : T->N;
        load_values64(c, p, vals);
        tri_exec<double, double, double>(c, p, TRI_SUM, alpha, b, x);
        if (c->opt.sync_checks) check_wait_error(c);
    });
}

// FP32 values and vectors, FP32 arithmetic in the same term order
int hala_trsv_f32(hala_tri T, char trans, float alpha, const float *vals, const float *b, float *x) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(T && vals && b && x, HALA_ERR_ARG, "hala_trsv_f32: null argument");
        hala_ctx c = T->ctx;
        set_device(c);
        bool tr = !(trans == 'N' || trans == 'n');
        HALA_REQUIRE(!tr || trans == 'T' || trans == 't' || trans == 'C' || trans == 'c', HALA_ERR_ARG,
                     "hala_trsv_f32: trans must be N, T or C");
        tri_plan &p = tr ? T->T : T->N;
        {  // the plan's loader takes FP64 values; the widening and its narrowing back are exact
            tmp_scope ts(c);
            dbuf &w = ts.get();
            w.reserve((size_t)T->nnz * sizeof(double));
            f32_to_f64<<<(int)((T->nnz + 255) / 256), 256, 0, c->stream>>>(T->nnz, vals, w.as<double>());
            HALA_CHECK_LAUNCH();
            load_values32(c, p, w.as<double>());
            HALA_HIP(hipStreamSynchronize(c->stream));
        }
        tri_exec<float, float, float>(c, p, TRI_SUM, alpha, b, x);
        if (c->opt.sync_checks) check_wait_error(c);
    });
}

int hala_trsm_f64(hala_tri T, char transa, char transb, int nrhs, double alpha, const double *vals, double *B,
                  int ldb) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(T && vals, HALA_ERR_ARG, "hala_trsm_f64: null argument");
        set_device(T->ctx);
        tri_trsm_cols(T, transb, nrhs, sizeof(double), B, ldb, [&](const void *in, void *out) {
            const int rc = hala_trsv_f64(T, transa, alpha, vals, (const double *)in, (double *)out);
            if (rc != HALA_OK) throw status_error{rc, hala_last_error()};
        });
    });
}

int hala_trsm_f32(hala_tri T, char transa, char transb, int nrhs, float alpha, const float *vals, float *B, int ldb) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(T && vals, HALA_ERR_ARG, "hala_trsm_f32: null argument");
        set_device(T->ctx);
        tri_trsm_cols(T, transb, nrhs, sizeof(float), B, ldb, [&](const void *in, void *out) {
            const int rc = hala_trsv_f32(T, transa, alpha, vals, (const float *)in, (float *)out);
            if (rc != HALA_OK) throw status_error{rc, hala_last_error()};
        });
    });
}

}  // extern "C"
