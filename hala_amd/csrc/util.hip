// util.hip -- scan, stable bucket sort, CSR transpose pattern.
#include <climits>

#include <hipcub/hipcub.hpp>

#include "util.hpp"

namespace hala_amd {

// out[0] = 0 and out[1..n] = the inclusive sums (hipCUB's decoupled look-back
// scan): the exclusive sums with the total in out[n].  (A three-pass scan
// with a single-block middle pass ran the 14.6 M pair counts of a 128^3 ILU
// at 0.35 TB/s.)
void exclusive_scan_i32(hala_ctx c, const int *in, int *out, int n) {
    HALA_REQUIRE(n >= 0, HALA_ERR_ARG, "scan: negative size");
    HALA_HIP(hipMemsetAsync(out, 0, sizeof(int), c->stream));
    if (n == 0) return;
    tmp_scope ts(c);
    dbuf &tmp = ts.get();
    size_t tbytes = 0;
    HALA_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tbytes, in, out + 1, n, c->stream));
    tmp.reserve(tbytes > 0 ? tbytes : 1);
    HALA_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tbytes, in, out + 1, n, c->stream));
}

// Items 0..n-1 (the sort's values).
__global__ void iota_i32(int n, int *__restrict__ v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

// bptr[b] = first position of bucket b in the sorted keys (a lower bound by
// bisection; empty buckets get the next bucket's start), bptr[nbuckets] = n.
__global__ void bucket_starts(int nbuckets, const int *__restrict__ keys, int n, int *__restrict__ bptr) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nbuckets) return;
    if (b == nbuckets) {
        bptr[b] = n;
        return;
    }
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < b) lo = mid + 1;
        else hi = mid;
    }
    bptr[b] = lo;
}

// A stable LSD radix sort of (bucket, item) pairs over the bucket's bits
// (hipCUB), so the items of a bucket come out ascending; the bucket offsets
// by bisection.  (Counting with one global atomic per item queued on the
// few hundred level counters: 1-1.4 ms per 128^3 triangle for the two sorts.)
void bucket_sort(hala_ctx c, int nitems, const int *bucket, int nbuckets, int *bptr, int *items) {
    HALA_REQUIRE(nitems >= 0 && nbuckets >= 1, HALA_ERR_ARG, "bucket_sort: bad sizes");
    if (nitems == 0) {
        HALA_HIP(hipMemsetAsync(bptr, 0, (size_t)(nbuckets + 1) * sizeof(int), c->stream));
        return;
    }
    int bits = 1;
    while (bits < 31 && (1 << bits) < nbuckets) bits++;
    tmp_scope ts(c);
    dbuf &keys = ts.get(), &vals = ts.get(), &tmp = ts.get();
    keys.reserve((size_t)nitems * sizeof(int));
    vals.reserve((size_t)nitems * sizeof(int));
    iota_i32<<<(nitems + 255) / 256, 256, 0, c->stream>>>(nitems, vals.as<int>());
    HALA_CHECK_LAUNCH();
    size_t tbytes = 0;
    HALA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, bucket, keys.as<int>(), vals.as<int>(), items,
                                                 nitems, 0, bits, c->stream));
    tmp.reserve(tbytes > 0 ? tbytes : 1);
    HALA_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tbytes, bucket, keys.as<int>(), vals.as<int>(), items,
                                                 nitems, 0, bits, c->stream));
    bucket_starts<<<(nbuckets + 256) / 256, 256, 0, c->stream>>>(nbuckets, keys.as<int>(), nitems, bptr);
    HALA_CHECK_LAUNCH();
}

__global__ void expand_rows_kernel(int nrows, const int *__restrict__ pntr, int *__restrict__ row_of) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    for (int j = pntr[i]; j < pntr[i + 1]; j++) row_of[j] = i;
}

void expand_rows(hala_ctx c, int nrows, const int *pntr, int *row_of) {
    if (nrows <= 0) return;
    expand_rows_kernel<<<(nrows + 255) / 256, 256, 0, c->stream>>>(nrows, pntr, row_of);
    HALA_CHECK_LAUNCH();
}

__global__ void gather_i32(int n, const int *__restrict__ idx, const int *__restrict__ src, int *__restrict__ dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

void build_transpose(hala_ctx c, int nrows, int ncols, int64_t nnz, const int *pntr, const int *indx,
                     csr_transpose &t) {
    HALA_REQUIRE(nnz < INT_MAX, HALA_ERR_ARG, "transpose: too many entries");
    t.ncols = ncols;
    t.nnz = nnz;
    t.tpntr.reserve((size_t)(ncols + 1) * sizeof(int));
    t.trow.reserve((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    t.tmap.reserve((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    // bucket = column of each entry; items = entry positions, ascending per column
    bucket_sort(c, (int)nnz, indx, ncols, t.tpntr.as<int>(), t.tmap.as<int>());
    if (nnz > 0) {
        dbuf row_of;
        row_of.reserve((size_t)nnz * sizeof(int));
        expand_rows(c, nrows, pntr, row_of.as<int>());
        gather_i32<<<(int)((nnz + 255) / 256), 256, 0, c->stream>>>((int)nnz, t.tmap.as<int>(), row_of.as<int>(),
                                                                    t.trow.as<int>());
        HALA_CHECK_LAUNCH();
        HALA_HIP(hipStreamSynchronize(c->stream));
    }
    t.built = true;
}

__global__ __launch_bounds__(256) void max_kernel(const int *__restrict__ x, int n, int *out) {
    __shared__ int sh[256];
    int m = INT_MIN;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = max(m, x[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, sh[0]);
}

__global__ void set_i32(int *p, int v) { *p = v; }

int reduce_max_i32(hala_ctx c, const int *x, int n) {
    int *out = reinterpret_cast<int *>(c->scal.as<double>() + S_DOT + 1);
    set_i32<<<1, 1, 0, c->stream>>>(out, INT_MIN);
    HALA_CHECK_LAUNCH();
    if (n > 0) {
        max_kernel<<<grid_for(n, 256 * 8, 1024), 256, 0, c->stream>>>(x, n, out);
        HALA_CHECK_LAUNCH();
    }
    return fetch_i32(c, out);
}

int fetch_i32(hala_ctx c, const int *p) {
    int v = 0;
    HALA_HIP(hipMemcpyAsync(c->host_pinned, p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HALA_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(&v, c->host_pinned, sizeof(int));
    return v;
}

}  // namespace hala_amd
