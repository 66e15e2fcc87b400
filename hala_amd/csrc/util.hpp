// util.hpp -- device-side building blocks shared by the analysis code.
#pragma once
#include "common.hpp"

namespace hala_amd {

// out[0..n]: exclusive prefix sum of in[0..n), out[n] = total (in and out may not alias).
void exclusive_scan_i32(hala_ctx c, const int *in, int *out, int n);

// Groups items 0..nitems-1 by bucket[item] (0 <= bucket < nbuckets):
// bptr[nbuckets+1] bucket offsets, items[] listed bucket by bucket, ascending
// inside each bucket (a stable sort).
void bucket_sort(hala_ctx c, int nitems, const int *bucket, int nbuckets, int *bptr, int *items);

// Transposed pattern of an nrows x ncols CSR: tpntr[ncols+1]; for each
// transposed entry e: trow[e] = source row, tmap[e] = source position;
// entries of a transposed row ascend in source position (= source row).
struct csr_transpose {
    int ncols = 0;
    int64_t nnz = 0;
    dbuf tpntr, trow, tmap;
    bool built = false;
};
void build_transpose(hala_ctx c, int nrows, int ncols, int64_t nnz, const int *pntr, const int *indx,
                     csr_transpose &t);

// row_of[j] = row owning entry j
void expand_rows(hala_ctx c, int nrows, const int *pntr, int *row_of);

// max over an int array (device), returned on the host (synchronises)
int reduce_max_i32(hala_ctx c, const int *x, int n);

// copy one int from the device (synchronises)
int fetch_i32(hala_ctx c, const int *p);

}  // namespace hala_amd
