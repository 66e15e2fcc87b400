// util.hip -- scan, stable bucket sort, CSR transpose pattern.
#include <climits>

#include "util.hpp"

namespace hala_amd {

constexpr int SCAN_B = 1024;  // elements per scan block (256 threads x 4)

__global__ __launch_bounds__(256) void scan_block_sums(const int *__restrict__ in, int n, int *__restrict__ sums) {
    __shared__ int sh[256];
    int base = blockIdx.x * SCAN_B + threadIdx.x * 4;
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (base + k < n) s += in[base + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = sh[0];
}

// In-place exclusive scan of sums[0..m) by one block; total into sums[m].
__global__ __launch_bounds__(1024) void scan_sums_single(int *sums, int m) {
    __shared__ int sh[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < m; base += 1024) {
        int i = base + threadIdx.x;
        int v = (i < m) ? sums[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive
            int t = (threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < m) sums[i] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[m] = carry;
}

__global__ __launch_bounds__(256) void scan_apply(const int *__restrict__ in, int n, const int *__restrict__ sums,
                                                  int *__restrict__ out) {
    __shared__ int sh[256];
    int base = blockIdx.x * SCAN_B + threadIdx.x * 4;
    int v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int t = (threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int run = sums[blockIdx.x] + sh[threadIdx.x] - s;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = sums[gridDim.x];
}

void exclusive_scan_i32(hala_ctx c, const int *in, int *out, int n) {
    HALA_REQUIRE(n >= 0, HALA_ERR_ARG, "scan: negative size");
    if (n == 0) {
        HALA_HIP(hipMemsetAsync(out, 0, sizeof(int), c->stream));
        return;
    }
    int nb = (n + SCAN_B - 1) / SCAN_B;
    tmp_scope ts(c);
    dbuf &sums = ts.get();
    sums.reserve((size_t)(nb + 1) * sizeof(int));
    scan_block_sums<<<nb, 256, 0, c->stream>>>(in, n, sums.as<int>());
    HALA_CHECK_LAUNCH();
    scan_sums_single<<<1, 1024, 0, c->stream>>>(sums.as<int>(), nb);
    HALA_CHECK_LAUNCH();
    scan_apply<<<nb, 256, 0, c->stream>>>(in, n, sums.as<int>(), out);
    HALA_CHECK_LAUNCH();
}

__global__ void count_buckets(int nitems, const int *__restrict__ bucket, int *__restrict__ cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nitems) atomicAdd(cnt + bucket[i], 1);
}

__global__ void scatter_buckets(int nitems, const int *__restrict__ bucket, int *__restrict__ cursor,
                                int *__restrict__ items) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nitems) {
        int p = atomicAdd(cursor + bucket[i], 1);
        items[p] = i;
    }
}

// Buckets of at most SMALL items: one thread, insertion sort in place.
constexpr int SMALL = 32;
__global__ void sort_small_buckets(int nbuckets, const int *__restrict__ bptr, int *__restrict__ items) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbuckets) return;
    int s = bptr[b], e = bptr[b + 1];
    if (e - s > SMALL) return;
    for (int i = s + 1; i < e; i++) {
        int v = items[i];
        int j = i - 1;
        while (j >= s && items[j] > v) {
            items[j + 1] = items[j];
            j--;
        }
        items[j + 1] = v;
    }
}

// Larger buckets: one 256-thread block, bitonic sort in LDS (up to LDS_SORT
// items), or bitonic sort in global memory with the same block beyond that.
constexpr int LDS_SORT = 8192;
__global__ __launch_bounds__(256) void sort_large_buckets(int nbuckets, const int *__restrict__ bptr,
                                                          int *__restrict__ items, int *__restrict__ scratch) {
    __shared__ int sh[LDS_SORT];
    int b = blockIdx.x;
    if (b >= nbuckets) return;
    int s = bptr[b], e = bptr[b + 1];
    int len = e - s;
    if (len <= SMALL) return;
    int p2 = 1;
    while (p2 < len) p2 <<= 1;
    int *buf;
    if (p2 <= LDS_SORT) {
        buf = sh;
    } else {
        buf = scratch + s * 2;  // scratch holds 2*nitems ints: room for p2 <= 2*len
    }
    for (int i = threadIdx.x; i < p2; i += 256) buf[i] = (i < len) ? items[s + i] : INT_MAX;
    __syncthreads();
    for (int k = 2; k <= p2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < p2; i += 256) {
                int l = i ^ j;
                if (l > i) {
                    int a = buf[i], c = buf[l];
                    bool up = ((i & k) == 0);
                    if ((a > c) == up) {
                        buf[i] = c;
                        buf[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < len; i += 256) items[s + i] = buf[i];
}

void bucket_sort(hala_ctx c, int nitems, const int *bucket, int nbuckets, int *bptr, int *items, bool sorted) {
    HALA_REQUIRE(nitems >= 0 && nbuckets >= 1, HALA_ERR_ARG, "bucket_sort: bad sizes");
    tmp_scope ts(c);
    dbuf &cnt = ts.get(), &cur = ts.get();
    cnt.reserve((size_t)(nbuckets + 1) * sizeof(int));
    cur.reserve((size_t)(nbuckets + 1) * sizeof(int));
    HALA_HIP(hipMemsetAsync(cnt.p, 0, (size_t)(nbuckets + 1) * sizeof(int), c->stream));
    if (nitems > 0) {
        count_buckets<<<(nitems + 255) / 256, 256, 0, c->stream>>>(nitems, bucket, cnt.as<int>());
        HALA_CHECK_LAUNCH();
    }
    exclusive_scan_i32(c, cnt.as<int>(), bptr, nbuckets);
    if (nitems == 0) return;
    HALA_HIP(hipMemcpyAsync(cur.p, bptr, (size_t)nbuckets * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    scatter_buckets<<<(nitems + 255) / 256, 256, 0, c->stream>>>(nitems, bucket, cur.as<int>(), items);
    HALA_CHECK_LAUNCH();
    if (!sorted) return;  // items of a bucket in scatter (arbitrary) order
    sort_small_buckets<<<(nbuckets + 255) / 256, 256, 0, c->stream>>>(nbuckets, bptr, items);
    HALA_CHECK_LAUNCH();
    // the largest bucket (sizes are cnt[]) decides whether the block-per-bucket pass is needed
    int maxlen = reduce_max_i32(c, cnt.as<int>(), nbuckets);
    if (maxlen > SMALL) {
        dbuf &scratch = ts.get();
        int p2 = 1;
        while (p2 < maxlen) p2 <<= 1;
        if (p2 > LDS_SORT) scratch.reserve((size_t)2 * nitems * sizeof(int) + (size_t)p2 * sizeof(int));
        sort_large_buckets<<<nbuckets, 256, 0, c->stream>>>(nbuckets, bptr, items,
                                                            p2 > LDS_SORT ? scratch.as<int>() : nullptr);
        HALA_CHECK_LAUNCH();
    }
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
