// spmv.hpp -- CSR matrix handle and the sparse matrix-vector product cores.
#pragma once
#include <functional>
#include "util.hpp"

struct hala_csr_s {
    hala_ctx ctx = nullptr;
    int nrows = 0, ncols = 0;
    int64_t nnz = 0;
    const int *pntr = nullptr;  // device, not owned
    const int *indx = nullptr;  // device, not owned
    int max_row = 0, min_row = 0;
    double mean_row = 0.0;
    int kernel = 0;  // 0: LDS-staged block of rows, 1: wavefront per row
    hala_amd::csr_transpose tp;  // built on first transposed product
    // column of every entry as an int16 offset from its 256-row block's first
    // row, when every offset fits (banded patterns: every stencil here); null
    // otherwise.  The LDS-staged products stream it instead of indx.
    hala_amd::dbuf boff;
};

namespace hala_amd {

constexpr int SPMV_ROWS = 256;  // rows per block in the LDS-staged kernel (= canonical block)

// Matrices whose values + indices reach this size are streamed with
// non-temporal loads: inside a solver they are evicted between products
// anyway (measured on the 217 MB 128^3 7-point operator from a cold MALL: 59 -> 50 us),
// while small ones stay cached across products.
constexpr int64_t NT_STREAM_BYTES = 64ll << 20;
template <class TV>
inline bool stream_nt(int64_t nnz) { return nnz * (int64_t)(sizeof(TV) + sizeof(int)) >= NT_STREAM_BYTES; }

template <bool NT, class T>
__device__ __forceinline__ T ld_stream(const T *p) {
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <class TA>
struct spmv_tile {
    static constexpr int value = sizeof(TA) == 8 ? 2048 : 4096;  // products staged in LDS per pass
};

// LDS-staged product of one block of SPMV_ROWS consecutive rows.  The block's
// nonzeros are streamed in coalesced tiles: every thread forms products
// vals[j] * x[indx[j]] into LDS, then each thread adds the products of its own
// row in CSR order.  Returns the row sum (TA(0) for rows past the end).
// SUB: start from `init` and subtract the products one at a time instead.
// NT: values and indices by non-temporal loads.
// TI = short: indx holds each entry's column minus the block's first row
// (hala_csr_s::boff), two bytes instead of four per entry.
// Must be called by all 256 threads of the block.
template <class TV, class TX, class TA, bool SUB = false, bool NT = false, class TI = int>
__device__ __forceinline__ TA csr_block_rowsum(int nrows, const int *__restrict__ pntr, const TI *__restrict__ indx,
                                               const TV *__restrict__ vals, const TX *__restrict__ x, TA *prod,
                                               int blk, TA init = TA(0)) {
    static_assert(sizeof(TI) == 4 || sizeof(TI) == 2, "int or short column indices");
    constexpr int TILE = spmv_tile<TA>::value;
    const int r0 = blk * SPMV_ROWS;
    const int cb = sizeof(TI) == 2 ? r0 : 0;  // column base of the stored indices
    const int row = r0 + threadIdx.x;
    const int rlast = min(r0 + SPMV_ROWS, nrows);
    const int bstart = pntr[r0];
    const int bend = pntr[rlast];
    int rs = 0, re = 0;
    if (row < nrows) {
        rs = pntr[row];
        re = pntr[row + 1];
    }
    TA acc = init;
    for (int base = bstart; base < bend; base += TILE) {
        const int cnt = min(TILE, bend - base);
        int e = threadIdx.x;
        for (; e + 3 * SPMV_ROWS < cnt; e += 4 * SPMV_ROWS) {
            const int j = base + e;
            int c0 = cb + (int)ld_stream<NT>(indx + j), c1 = cb + (int)ld_stream<NT>(indx + j + SPMV_ROWS),
                c2 = cb + (int)ld_stream<NT>(indx + j + 2 * SPMV_ROWS),
                c3 = cb + (int)ld_stream<NT>(indx + j + 3 * SPMV_ROWS);
            TV v0 = ld_stream<NT>(vals + j), v1 = ld_stream<NT>(vals + j + SPMV_ROWS),
               v2 = ld_stream<NT>(vals + j + 2 * SPMV_ROWS), v3 = ld_stream<NT>(vals + j + 3 * SPMV_ROWS);
            TA x0 = (TA)x[c0], x1 = (TA)x[c1], x2 = (TA)x[c2], x3 = (TA)x[c3];
            prod[e] = (TA)v0 * x0;
            prod[e + SPMV_ROWS] = (TA)v1 * x1;
            prod[e + 2 * SPMV_ROWS] = (TA)v2 * x2;
            prod[e + 3 * SPMV_ROWS] = (TA)v3 * x3;
        }
        for (; e < cnt; e += SPMV_ROWS) {
            const int j = base + e;
            prod[e] = (TA)ld_stream<NT>(vals + j) * (TA)x[cb + (int)ld_stream<NT>(indx + j)];
        }
        __syncthreads();
        const int a = max(rs, base), b = min(re, base + cnt);
        for (int j = a; j < b; j++) acc = SUB ? acc - prod[j - base] : acc + prod[j - base];
        __syncthreads();
    }
    return acc;
}

template <class TV, class TX, class TA>
__device__ __forceinline__ TA csr_block_rowsum(int nrows, const int *__restrict__ pntr, const int *__restrict__ indx,
                                               const TV *__restrict__ vals, const TX *__restrict__ x, TA *prod) {
    return csr_block_rowsum<TV, TX, TA>(nrows, pntr, indx, vals, x, prod, blockIdx.x);
}

// The int16 copy of a pattern's columns as offsets from the first row of each
// entry's 256-row block (what csr_block_rowsum<..., short> reads), built into
// out when every offset fits; false (out untouched) otherwise.  Reads two
// counters back (synchronises the stream).
bool build_block_offsets(hala_ctx c, int nrows, int64_t nnz, const int *ptr, const int *idx, dbuf &out);

// First row of an invalid pattern (-1 when valid): pointers must start at 0,
// end at nnz and not decrease; columns sorted, unique, inside [0, ncols);
// require_diag: every row holds its diagonal.
int validate_pattern(hala_ctx c, int nrows, int ncols, int64_t nnz, const int *pntr, const int *indx,
                     int require_diag);

// Launch helpers (defined in spmv.hip)
template <class TV, class TX, class TY>
void spmv_axpby(hala_csr A, const TV *vals, const TX *x, TY alpha, TY beta, TY *y, bool f32_inner);
template <class T>
void spmv_trans(hala_csr A, T alpha, const T *vals, const T *x, T beta, T *y);
void ensure_transpose(hala_csr A);

// n elements of esize (4, 8 or 16) bytes from src (stride ss elements) to dst (stride ds), on the context's stream
void strided_copy(hala_ctx c, int n, size_t esize, const void *src, int64_t ss, void *dst, int64_t ds);

// Sparse matrix times dense matrix, column by column as the CPU engine orders it
// (sparse/hala_sparse_utils.hpp:127-167): C[:, i] = alpha op(A) op(B)[:, i] + beta C[:, i]
// through `mv` (one product from a contiguous input x of op(A)'s column count
// into y).  op(B) column i is B's column i ('N') or row i ('T', 'C'; `conj`
// conjugates the gathered copy for 'C').
void spmm_cols(hala_csr A, char transa, char transb, int b_rows, int b_cols, size_t esize, const void *B, int ldb,
               void *C, int ldc, const std::function<void(const void *, void *)> &mv,
               const std::function<void(void *, int)> &conj);

}  // namespace hala_amd
