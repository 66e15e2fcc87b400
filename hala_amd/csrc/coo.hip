// coo.hip -- COO triplets to sorted CSR on the device: the ingestion step in
// front of the solvers, which (like the reference, sparse/hala_sparse_core.hpp:42
// and :56-59) need every row's columns ascending and unique and, for ILU(0),
// the diagonal present.  Entries are ordered by (row, column, input position);
// duplicates are summed left to right in input order; optionally a zero is
// inserted where a diagonal entry is missing.
#include <climits>

#include "util.hpp"

namespace hala_amd {

namespace {

__global__ void set_i32_k(int *p, int v) { *p = v; }

__global__ void coo_check(int64_t nent, int n, int m, const int *__restrict__ rows, const int *__restrict__ cols,
                          int *bad) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    const int r = rows[e], cl = cols[e];
    if (r < 0 || r >= n || cl < 0 || cl >= m) atomicMin(bad, (int)e);
}

// entries of every row (input order after the bucket sort) by column, stable
__global__ void coo_sort_rows(int n, const int *__restrict__ rptr, int *__restrict__ perm,
                              const int *__restrict__ cols) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int a = rptr[r], b = rptr[r + 1];
    for (int k = a + 1; k < b; k++) {
        const int e = perm[k], ce = cols[e];
        int q = k - 1;
        while (q >= a && cols[perm[q]] > ce) {
            perm[q + 1] = perm[q];
            q--;
        }
        perm[q + 1] = e;
    }
}

// distinct columns of every row (+ the diagonal when missing and requested)
__global__ void coo_count(int n, int m, int add_diag, const int *__restrict__ rptr, const int *__restrict__ perm,
                          const int *__restrict__ cols, int *__restrict__ cnt) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int u = 0, prev = -1;
    bool dg = false;
    for (int k = rptr[r]; k < rptr[r + 1]; k++) {
        const int cl = cols[perm[k]];
        if (cl != prev) u++;
        dg |= cl == r;
        prev = cl;
    }
    if (add_diag && !dg && r < m) u++;
    cnt[r] = u;
}

__global__ void coo_fill(int n, int m, int add_diag, const int *__restrict__ rptr, const int *__restrict__ perm,
                         const int *__restrict__ cols, const double *__restrict__ vals, const int *__restrict__ pntr,
                         int *__restrict__ indx, double *__restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int o = pntr[r];
    const int a = rptr[r], b = rptr[r + 1];
    bool need_diag = add_diag && r < m;
    for (int k = a; k < b;) {
        const int cl = cols[perm[k]];
        if (need_diag && cl > r) {
            indx[o] = r;
            out[o++] = 0.0;
            need_diag = false;
        }
        double s = vals[perm[k]];
        for (k++; k < b && cols[perm[k]] == cl; k++) s = s + vals[perm[k]];
        if (cl == r) need_diag = false;
        indx[o] = cl;
        out[o++] = s;
    }
    if (need_diag) {
        indx[o] = r;
        out[o] = 0.0;
    }
}

}  // namespace
}  // namespace hala_amd

using namespace hala_amd;

extern "C" int hala_coo_to_csr(hala_ctx c, int nrows, int ncols, int64_t nent, const int *rows, const int *cols,
                               const double *vals, int add_diag, int *pntr, int *indx, double *vals_out,
                               int64_t *nnz) {
    return guard([&] {
        HALA_REQUIRE(c && pntr && nnz && (nent == 0 || (rows && cols)), HALA_ERR_ARG,
                     "hala_coo_to_csr: null argument");
        HALA_REQUIRE(nrows >= 0 && ncols >= 0 && nent >= 0 && nent + nrows < INT_MAX, HALA_ERR_ARG,
                     "hala_coo_to_csr: bad sizes");
        HALA_REQUIRE(!indx || ((vals || nent == 0) && vals_out), HALA_ERR_ARG, "hala_coo_to_csr: values missing");
        set_device(c);
        const int n = nrows;
        dbuf bad, rptr, perm, cnt;
        bad.reserve(sizeof(int));
        set_i32_k<<<1, 1, 0, c->stream>>>(bad.as<int>(), INT_MAX);
        if (nent > 0) {
            coo_check<<<(int)((nent + 255) / 256), 256, 0, c->stream>>>(nent, nrows, ncols, rows, cols, bad.as<int>());
            HALA_CHECK_LAUNCH();
        }
        const int b = fetch_i32(c, bad.as<int>());
        HALA_REQUIRE(b == INT_MAX, HALA_ERR_PATTERN,
                     "hala_coo_to_csr: entry " + std::to_string(b) + " has a row or column out of range");
        rptr.reserve((size_t)(n + 1) * sizeof(int));
        perm.reserve((size_t)(nent > 0 ? nent : 1) * sizeof(int));
        cnt.reserve((size_t)(n + 1) * sizeof(int));
        if (n == 0) {
            set_i32_k<<<1, 1, 0, c->stream>>>(pntr, 0);
            *nnz = 0;
            HALA_HIP(hipStreamSynchronize(c->stream));
            return;
        }
        bucket_sort(c, (int)nent, rows, n, rptr.as<int>(), perm.as<int>());
        const int g = (n + 255) / 256;
        coo_sort_rows<<<g, 256, 0, c->stream>>>(n, rptr.as<int>(), perm.as<int>(), cols);
        coo_count<<<g, 256, 0, c->stream>>>(n, ncols, add_diag ? 1 : 0, rptr.as<int>(), perm.as<int>(), cols,
                                            cnt.as<int>());
        HALA_CHECK_LAUNCH();
        exclusive_scan_i32(c, cnt.as<int>(), pntr, n);
        *nnz = fetch_i32(c, pntr + n);
        if (indx) {
            coo_fill<<<g, 256, 0, c->stream>>>(n, ncols, add_diag ? 1 : 0, rptr.as<int>(), perm.as<int>(), cols, vals,
                                               pntr, indx, vals_out);
            HALA_CHECK_LAUNCH();
        }
        HALA_HIP(hipStreamSynchronize(c->stream));
    });
}
