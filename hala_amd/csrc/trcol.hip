// trcol.hip -- column-blocked execution of the level-scheduled triangular sweeps.
//
// The rows are cut into `ncol` blocks ("columns").  Plainly, contiguous row
// ranges: every dependency of a forward sweep points to a lower row and of a
// backward sweep to a higher row, so it lies in the same column or in one
// handled before it.  For a stencil-like pattern whose rows come in "planes"
// (P rows, the middle of the largest dependency offsets) a column is instead
// one of B pieces of a slab of planes: the column a row's dependencies lie in
// is then the same, the previous piece or the previous slab, and a sweep's
// longest chain of columns shrinks from A*B to A+B-1 (each link costs a
// global round trip before the next column can start).  Any assignment is
// correct, since columns only exchange values through the polled output.  One
// workgroup owns a column and walks the levels in order; the rows of its
// column at one level (a "cell", a contiguous piece of that level's row list)
// get one thread each.  A dependency produced by the same workgroup within the
// last K levels is read from an LDS window; every other dependency ("external")
// is polled in the level-ordered output with agent-scope loads issued several
// levels ahead and staged into LDS just before its level.  A level ends with an
// LDS-only wait and a workgroup barrier, so the critical path pays a global
// round trip only where it crosses to another column, instead of at every
// level.  Per row the terms are applied in the row's own order, so results
// are bit-identical to the level-pool sweep and to the sequential reference.
#include <algorithm>
#include <climits>

#include "trsolve.hpp"

namespace hala_amd {

namespace {

constexpr int CI = 5;       // ints per (column, level) cell record
constexpr int KMAX = 8;     // levels held in the LDS window at most
constexpr int BTMAX = 1024;  // padding of slot-indexed arrays (idle-lane reads)
constexpr int CPMAX = 512;   // rows of a cell at most (= compute threads of a column workgroup)
constexpr int CPMAX_LONG = 256;  // the same for long-row records (register budget of their prefetch)
constexpr int STEP_BYTES = 20 * 1024;  // automatic column count: record bytes one level of a column reads at most
// the same for columns split across planes: they pass fewer round trips per
// sweep, so narrower cells pay (128^3 7-point apply: 0.53 ms at 256 rows per
// cell, 0.39 ms at 128, 0.41 ms at 64; contiguous columns 0.58 ms)
constexpr int STEP_BYTES_SPLIT = 8 * 1024;
constexpr int ELMAX = 8;     // external values per loader lane at most (512 per cell)
constexpr int DEDUP_MAX = 4096;  // external references of one cell at most
// Levels an external value is polled ahead of its use: 2 for short-row
// operators (128^3 7-point apply 0.577 vs 0.596 ms), 1 for long rows (whose columns
// already lag their producers by several levels: 1.33 vs 1.54 ms at 2).
constexpr int PE_MAX = 2;
constexpr int poll_ahead(int maxd) { return maxd == 4 ? 2 : 1; }
constexpr int64_t LDS_BUDGET = 64 * 1024;

// Column of every row.  Rows are taken in sweep order (r' = n-1-r backward).
// B == 1: A contiguous blocks.  B > 1: slab a = plane * A / (n / P) of whole
// planes of P rows, piece b = (r' mod P) * B / P, column a * B + b.
__global__ void col_map(int n, int backward, int A, int B, int P, int *__restrict__ cmap) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t k = backward ? (int64_t)(n - 1 - r) : (int64_t)r;
    if (B == 1) {
        cmap[r] = (int)(k * A / n);
    } else {
        const int64_t nz = n / P, plane = k / P, pos = k % P;
        cmap[r] = (int)(plane * A / nz) * B + (int)(pos * B / P);
    }
}

// per row: the largest dependency offset |r - d|, or (lo > 0) minus the
// smallest offset above lo (INT_MIN + 1 for none) -- max-reduced afterwards
__global__ void dep_offsets(int n, int lo, const int *__restrict__ dptr, const int *__restrict__ ddep,
                            int *__restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int mx = 0, mn = INT_MAX;
    for (int e = dptr[r]; e < dptr[r + 1]; e++) {
        const int o = abs(r - ddep[e]);
        mx = max(mx, o);
        if (o > lo) mn = min(mn, o);
    }
    out[r] = lo > 0 ? (mn == INT_MAX ? INT_MIN + 1 : -mn) : mx;
}

// flag: some dependency lies in a higher-numbered column
__global__ void col_upward(int n, const int *__restrict__ cmap, const int *__restrict__ dptr,
                           const int *__restrict__ ddep, int *__restrict__ flag) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int c = cmap[r];
    bool up = false;
    for (int e = dptr[r]; e < dptr[r + 1]; e++) up |= cmap[ddep[e]] > c;
    // one store per wavefront, and none once the flag is seen set (stores of
    // every row to one word queue on its memory channel, and other kernels'
    // accesses behind them)
    if (__any(up) && (threadIdx.x & (WAVE - 1)) == 0 && ld_agent(flag) == 0) *flag = 1;
}

// the largest level gap of a same-column dependency (reduced over the
// wavefront first, then an atomic only where it would raise the counter's
// current value: the counter settles within the first wavefronts, and
// atomics of every wavefront on one word would queue on its memory channel)
__global__ __launch_bounds__(256) void col_gap(int n, const int *__restrict__ cmap, const int *__restrict__ lev,
                                               const int *__restrict__ dptr, const int *__restrict__ ddep,
                                               int *__restrict__ maxgap) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    int g = 0;
    if (r < n) {
        const int c = cmap[r];
        for (int e = dptr[r]; e < dptr[r + 1]; e++) {
            int d = ddep[e];
            if (cmap[d] == c) g = max(g, lev[r] - lev[d]);
        }
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) g = max(g, __shfl_xor(g, o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0 && g > 0 && g > ld_agent(maxgap)) atomicMax(maxgap, g);
}

// cell populations and the start of every cell inside its level's row list,
// from the bucket starts of the rows sorted by (level, column, row)
__global__ void col_cellcnt(int64_t ncell, int ncol, const int *__restrict__ bptr, int *__restrict__ cellcnt,
                            int *__restrict__ cstart) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncell) return;
    cellcnt[k] = bptr[k + 1] - bptr[k];
    cstart[k] = bptr[k] - bptr[k / ncol * ncol];
}

__device__ __forceinline__ bool is_local(int r, int d, int c, int cd, const int *lev, int K) {
    return cd == c && lev[r] - lev[d] < K;
}

// external dependencies per slot (ecnt zeroed beforehand: slots without a
// row keep 0).  A thread per row, in row order: the row's data and its
// neighbours' are read nearly contiguously, only the count lands scattered.
__global__ void col_extcnt(int n, const int *__restrict__ cmap, int K, const int *__restrict__ slot_of_row,
                           const int *__restrict__ lev, const int *__restrict__ dptr, const int *__restrict__ ddep,
                           int *__restrict__ ecnt) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int c = cmap[r];
    int k = 0;
    for (int e = dptr[r]; e < dptr[r + 1]; e++) {
        int d = ddep[e];
        if (!is_local(r, d, c, cmap[d], lev, K)) k++;
    }
    ecnt[slot_of_row[r]] = k;
}

// cell records (column-major) and the level range of every column
__global__ void col_cells(int nlev, int ncol, const int *__restrict__ lchunk, const int *__restrict__ cellcnt,
                          const int *__restrict__ cstart, const int *__restrict__ epos, int *__restrict__ ci,
                          int *__restrict__ colrange, int *__restrict__ maxext) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)nlev * ncol) return;
    const int l = (int)(id / ncol), c = (int)(id % ncol);
    const int S0 = lchunk[l] * WAVE, Wl = (lchunk[l + 1] - lchunk[l]) * WAVE;
    const int q0 = cstart[id], size = cellcnt[id];
    const int slot0 = S0 + q0;
    const int e0 = epos[slot0], en = epos[slot0 + size] - e0;
    int *o = ci + ((int64_t)c * nlev + l) * CI;
    o[0] = slot0;
    o[1] = q0;
    o[2] = size | (en << 16);
    o[3] = e0;
    o[4] = Wl;
    if (size > 0) {
        atomicMin(colrange + 2 * c, l);
        atomicMax(colrange + 2 * c + 1, l);
    }
    if (en > 0 && en > ld_agent(maxext)) atomicMax(maxext, en);  // (the atomic only where it raises the value)
}

// dependency codes: >= 0 LDS window index, -1 none, <= -2 index into the cell's external list
// (a thread per row, in row order, as col_extcnt; slots without a row keep the
// prefilled -1 codes)
__global__ void col_records(int n, const int *__restrict__ cmap, int ncol, int K, int CP, int L, int nlev,
                            const int *__restrict__ lev, const int *__restrict__ lchunk,
                            const int *__restrict__ cstart, const int *__restrict__ slot_of_row,
                            const int *__restrict__ dptr, const int *__restrict__ ddep, const int *__restrict__ dsrc,
                            const int *__restrict__ epos, const int *__restrict__ colrange, int *__restrict__ code,
                            int *__restrict__ src, int *__restrict__ ext_slot, int *__restrict__ spar) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int s = slot_of_row[r];
    const int l = lev[r];
    spar[s] = (l - colrange[2 * cmap[r]]) & 1;
    const int c = cmap[r];
    const int64_t base = (int64_t)s * L;  // thread-major: the row's L entries are contiguous
    const int a = dptr[r], len = dptr[r + 1] - a;
    int j = 0;
    for (int k = 0; k < L; k++) {
        int cd = -1, sr = -1;
        if (k < len) {
            const int d = ddep[a + k];
            const int dc = cmap[d];
            const int dl = lev[d];
            const int ds = slot_of_row[d];
            if (is_local(r, d, c, dc, lev, K)) {
                const int dS0 = lchunk[dl] * WAVE;
                cd = (dl % K) * CP + (ds - dS0 - cstart[(int64_t)dl * ncol + dc]);
            } else {
                ext_slot[epos[s] + j] = ds;
                cd = -2 - (epos[s] + j);  // raw reference; replaced by the cell's unique index below
                j++;
            }
            sr = dsrc[a + k];
        }
        code[base + k] = cd;
        src[base + k] = sr;
    }
}

// One workgroup per cell: sort the cell's external references, keep each
// slot once (written back to the front of the cell's segment) and record for
// every raw reference its index in that unique list.
// LDS: 3 arrays of cap ints (cap: a power of two >= the largest cell's raw
// count, at most DEDUP_MAX), so small lists do not cost occupancy.
__global__ __launch_bounds__(256) void col_dedup(int nlev, int ncol, int cap, int *__restrict__ ci,
                                                 int *__restrict__ ext_slot, int *__restrict__ uidx,
                                                 int *__restrict__ maxext, int *__restrict__ fail) {
    extern __shared__ int dsh[];
    int *key = dsh, *orig = dsh + cap, *flag = dsh + 2 * cap;
    __shared__ int nu_s;
    const int cell = blockIdx.x;
    const int l = cell / ncol, c = cell % ncol;
    int *rec = ci + ((int64_t)c * nlev + l) * CI;
    const int e0 = rec[3], en = rec[2] >> 16, size = rec[2] & 0xffff;
    if (en == 0) return;
    if (en > cap) {
        if (threadIdx.x == 0) atomicOr(fail, 1);
        return;
    }
    int np = 1;
    while (np < en) np <<= 1;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        const int v = i < en ? ext_slot[e0 + i] : INT_MAX;
        key[i] = v;
        if (i < en) orig[i] = v;
    }
    __syncthreads();
    for (int k = 2; k <= np; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < np; i += blockDim.x) {
                const int o = i ^ j;
                if (o > i) {
                    const int a = key[i], b = key[o];
                    const bool up = (i & k) == 0;
                    if (up ? (a > b) : (a < b)) {
                        key[i] = b;
                        key[o] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < en; i += blockDim.x) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
    __syncthreads();
    if (threadIdx.x == 0) {  // compaction in order (en is small)
        int u = 0;
        for (int i = 0; i < en; i++)
            if (flag[i]) key[u++] = key[i];
        nu_s = u;
    }
    __syncthreads();
    const int nu = nu_s;
    for (int i = threadIdx.x; i < nu; i += blockDim.x) ext_slot[e0 + i] = key[i];
    for (int i = threadIdx.x; i < en; i += blockDim.x) {
        const int v = orig[i];
        int lo = 0, hi = nu - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (key[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        uidx[e0 + i] = lo;
    }
    if (threadIdx.x == 0) {
        rec[2] = size | (nu << 16);
        if (nu > ld_agent(maxext)) atomicMax(maxext, nu);
    }
}

// raw external references in the codes -> the cell's unique index
__global__ void col_fixcodes(int64_t nrec, int *__restrict__ code, const int *__restrict__ uidx) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrec) return;
    const int cd = code[e];
    if (cd <= -2) code[e] = -2 - uidx[-cd - 2];
}

__global__ void row_lens(int n, const int *__restrict__ dptr, int *__restrict__ len) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) len[r] = dptr[r + 1] - dptr[r];
}

__global__ void init_range(int ncol, int nlev, int *__restrict__ colrange) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncol) {
        colrange[2 * c] = nlev;
        colrange[2 * c + 1] = -1;
    }
}

// Packed row records, one per slot, RB bytes each (a multiple of 16, read
// with 16-byte loads): int16 dependency codes [LR] | diagonal T | values T[LR]
// (zero padding after).  A thread per 16-byte quad of a record, so that each
// lane stores one whole quad next to its neighbours'.
template <class T>
__global__ void col_pack(int nslots, int nslots_pad, int LR, int RB, int kcp, int xc, const int *__restrict__ code,
                         const int *__restrict__ spar, const int *__restrict__ src, const int *__restrict__ dg_src,
                         const double *__restrict__ f, unsigned char *__restrict__ rec) {
    constexpr int SZ = (int)sizeof(T);
    const int nq = RB / 16;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nslots_pad * nq) return;
    const int s = (int)(t / nq), qd = (int)(t % nq);
    const bool in = s < nslots;
    const int zidx = kcp + 2 * xc;
    const int cb = 2 * LR, vb = cb + SZ, ve = vb + LR * SZ;  // ends of the codes, the diagonal, the values
    unsigned w4[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int b = 16 * qd + 4 * j;  // byte offset of the word in the record
        unsigned w = 0;
        if (b < cb) {
            // kernel-side codes are LDS element indices: a window entry, external e
            // of the slot's level in the buffer of that level's parity, or the zero
            // element zidx for no dependency (value 0)
            unsigned h[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int k = b / 2 + u;
                const int c = in ? code[(int64_t)s * LR + k] : -1;
                const int v = c == -1 ? zidx : (c >= 0 ? c : kcp + spar[s] * xc + (-c - 2));
                h[u] = (unsigned)(unsigned short)(short)v;
            }
            w = h[0] | (h[1] << 16);
        } else if (b < ve) {
            T val;
            int off;
            if (b < vb) {
                const int d = in ? dg_src[s] : -1;
                val = d >= 0 ? (T)f[d] : T(1);
                off = b - cb;
            } else {
                const int k = (b - vb) / SZ;
                const int c = in ? code[(int64_t)s * LR + k] : -1;
                const int e = (in && c != -1) ? src[(int64_t)s * LR + k] : -1;
                val = e >= 0 ? (T)f[e] : T(0);
                off = (b - vb) % SZ;
            }
            if (SZ == 8) {
                const unsigned long long bits = (unsigned long long)__double_as_longlong((double)val);
                w = off == 0 ? (unsigned)bits : (unsigned)(bits >> 32);
            } else {
                w = __float_as_uint((float)val);
            }
        }
        w4[j] = w;
    }
    reinterpret_cast<uint4 *>(rec)[t] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
}

// window + external double buffer + cell records (padded to whole pipeline blocks plus the read-ahead)
__global__ void fill_i32_k(int *__restrict__ x, int64_t n, int v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// widest level, in slots (64-row chunks)
int maxw_slots(const tri_plan &p) {
    int w = 0;
    for (int l = 0; l < p.nlev; l++) w = std::max(w, (p.h_level_ptr[l + 1] - p.h_level_ptr[l] + WAVE - 1) / WAVE * WAVE);
    return w;
}

tri_colplan &tri_colplan_reset(tri_colplan &q) {
    q.ok = false;
    q.code.release();
    q.src.release();
    q.ext_slot.release();
    q.spar.release();
    return q;
}

// window (K levels x CP rows) + external double buffer (2 x 64 EL values) +
// one zero element + slots of the polls in flight (up to PE_MAX x 64 EL) + cell records
int64_t lds_bytes(int K, int CP, int EL, int nlevcol, size_t tsize) {
    const int pad = nlevcol + 3 * 8 + 2;  // read-ahead of the deepest pipeline (P = 8)
    const int64_t xc = (int64_t)WAVE * EL;
    return (int64_t)K * CP * (int64_t)tsize + (2 * xc + 2) * (int64_t)tsize + PE_MAX * xc * 4 + (int64_t)pad * CI * 4;
}

// bytes of one packed row record: int16 codes [LR] | diagonal | values [LR], 16-byte multiple
int rec_bytes(int LR, size_t tsize) {
    const int b = 2 * LR + (int)tsize * (LR + 1);
    return (b + 15) / 16 * 16;
}

// loader-lane capacity for a cell with m distinct externals (0 when m is too large)
int ext_lanes(int m) {
    int el = 2;
    while (el * WAVE < m) el *= 2;
    return el <= ELMAX ? el : 0;
}

}  // namespace

bool build_col_plan(hala_ctx c, tri_plan &p, int ncol_req) {
    p.col.ok = false;
    p.col.why = 1;
    const int n = p.n, nlev = p.nlev, nslots = p.nchunk * WAVE;
    if (n < 2 * WAVE || nlev < 2) return false;
    int maxw = 0;
    for (int l = 0; l < nlev; l++) maxw = std::max(maxw, p.h_level_ptr[l + 1] - p.h_level_ptr[l]);
    tmp_scope ts(c);
    dbuf &lens = ts.get();
    lens.reserve((size_t)n * sizeof(int));
    row_lens<<<(n + 255) / 256, 256, 0, c->stream>>>(n, p.dptr.as<int>(), lens.as<int>());
    HALA_CHECK_LAUNCH();
    const int maxlen = reduce_max_i32(c, lens.as<int>(), n);
    p.col.why = 2;
    if (maxlen > MAXD || maxlen < 1) return false;
    const int g = (n + 255) / 256;
    dbuf &cellcnt = ts.get(), &gap = ts.get(), &cstart = ts.get(), &ecnt = ts.get(), &epos = ts.get();
    dbuf &maxext = ts.get(), &cmap = ts.get(), &cbptr = ts.get(), &corder = ts.get();
    dbuf ci, colrange;  // handed to the plan
    gap.reserve(sizeof(int));
    maxext.reserve(sizeof(int));
    cmap.reserve((size_t)n * sizeof(int));
    // record width: the kernel variant's dependency capacity (4 or 16)
    const int L = maxlen <= 4 ? 4 : MAXD;
    const int cpmax = L == 4 ? CPMAX : CPMAX_LONG;
    // plane size: the middle of the dependency offsets above half the largest
    int P = 0;
    if (c->opt.tri_col_split != 1) {
        dep_offsets<<<g, 256, 0, c->stream>>>(n, 0, p.dptr.as<int>(), p.ddep.as<int>(), lens.as<int>());
        HALA_CHECK_LAUNCH();
        const int dmax = reduce_max_i32(c, lens.as<int>(), n);
        if (dmax >= 2) {
            dep_offsets<<<g, 256, 0, c->stream>>>(n, dmax / 2, p.dptr.as<int>(), p.ddep.as<int>(), lens.as<int>());
            HALA_CHECK_LAUNCH();
            const int dmin = -reduce_max_i32(c, lens.as<int>(), n);
            P = (int)(((int64_t)dmax + dmin) / 2);
        }
        if (P <= 0 || n % P != 0) P = 0;
    }
    int ncol = ncol_req > 0 ? ncol_req : 4;
    while (ncol_req <= 0 && (int64_t)maxw > (int64_t)ncol * cpmax) ncol *= 2;
    for (; ncol <= 512; ncol *= 2) {
        if (ncol_req > 0 && ncol != ncol_req) break;
        // pieces per slab: the option, else the largest power of two B with B * B <= ncol
        int B = c->opt.tri_col_split > 0 ? c->opt.tri_col_split : 1;
        if (c->opt.tri_col_split == 0)
            while (4 * B * B <= ncol) B *= 2;
        if (B > 1 && (P == 0 || ncol % B != 0 || (int64_t)P < (int64_t)B * WAVE || n / P < ncol / B)) B = 1;
        if (B > 1) {
            // Contiguous columns only wait on lower-numbered ones.  A split one
            // may wait on a higher-numbered one (e.g. a 27-point row on the next
            // piece of the previous plane): two columns then trade values every
            // level (measured 1.6x slower on a 27-point operator), and every
            // column needs a CU of its own to be sure to progress.  Automatic
            // splits are therefore kept only where no dependency points upwards.
            col_map<<<g, 256, 0, c->stream>>>(n, p.backward ? 1 : 0, ncol / B, B, P, cmap.as<int>());
            HALA_CHECK_LAUNCH();
            HALA_HIP(hipMemsetAsync(gap.p, 0, sizeof(int), c->stream));
            col_upward<<<g, 256, 0, c->stream>>>(n, cmap.as<int>(), p.dptr.as<int>(), p.ddep.as<int>(),
                                                gap.as<int>());
            HALA_CHECK_LAUNCH();
            const bool upward = fetch_i32(c, gap.as<int>()) != 0;
            if (upward && (c->opt.tri_col_split == 0 || ncol > c->num_cus)) B = 1;
        }
        const int A = ncol / B;
        // the widest level spreads over ncol cells, so one holds at least maxw / ncol
        // of its rows: skip column counts whose cells cannot fit the budget
        if (ncol_req <= 0 && ncol < 512) {
            const int64_t cap = std::min<int64_t>(
                cpmax, (B > 1 ? STEP_BYTES_SPLIT : STEP_BYTES) / rec_bytes(L, sizeof(double)) / WAVE * WAVE);
            if ((int64_t)ncol * cap < maxw) continue;
        }
        if (B == 1) {
            col_map<<<g, 256, 0, c->stream>>>(n, p.backward ? 1 : 0, A, 1, 1, cmap.as<int>());
            HALA_CHECK_LAUNCH();
        }
        const int64_t ncell = (int64_t)nlev * ncol;
        HALA_REQUIRE(ncell < INT_MAX, HALA_ERR_ARG, "column plan: too many cells");
        cellcnt.reserve((size_t)ncell * sizeof(int));
        cstart.reserve((size_t)ncell * sizeof(int));
        cbptr.reserve((size_t)(ncell + 1) * sizeof(int));
        corder.reserve((size_t)n * sizeof(int));
        HALA_HIP(hipMemsetAsync(gap.p, 0, sizeof(int), c->stream));
        col_gap<<<g, 256, 0, c->stream>>>(n, cmap.as<int>(), p.lev.as<int>(), p.dptr.as<int>(), p.ddep.as<int>(),
                                         gap.as<int>());
        HALA_CHECK_LAUNCH();
        // rows by (level, column, row): the cells' populations and starts come
        // from the bucket starts, and the order is the layout's if this column
        // count is kept
        sort_level_keys(c, p, cmap.as<int>(), ncol, cbptr.as<int>(), corder.as<int>());
        col_cellcnt<<<(int)((ncell + 255) / 256), 256, 0, c->stream>>>(ncell, ncol, cbptr.as<int>(),
                                                                      cellcnt.as<int>(), cstart.as<int>());
        HALA_CHECK_LAUNCH();
        const int maxcell = reduce_max_i32(c, cellcnt.as<int>(), (int)ncell);
        const int CP = (maxcell + WAVE - 1) / WAVE * WAVE;
        p.col.ncol = ncol;
        p.col.CP = CP;
        p.col.why = 3;
        if (CP > cpmax) continue;
        // automatic choice: narrower cells until a level's records fit the per-step budget
        if (ncol_req <= 0 && ncol < 512 &&
            (int64_t)CP * rec_bytes(L, sizeof(double)) > (B > 1 ? STEP_BYTES_SPLIT : STEP_BYTES))
            continue;
        const int K = std::min(fetch_i32(c, gap.as<int>()) + 1, KMAX);
        // cells must be contiguous in every level's row list: rows of a level
        // ordered by (column, row), which also lays out the slots this layout
        // indexes (chunk_row, slot_of_row)
        apply_level_order(c, p, corder.as<int>());
        // external counts, cell starts, cell records
        ecnt.reserve((size_t)(nslots + 1) * sizeof(int));
        epos.reserve((size_t)(nslots + 1) * sizeof(int));
        HALA_HIP(hipMemsetAsync(ecnt.p, 0, (size_t)nslots * sizeof(int), c->stream));
        col_extcnt<<<g, 256, 0, c->stream>>>(n, cmap.as<int>(), K, p.slot_of_row.as<int>(), p.lev.as<int>(),
                                             p.dptr.as<int>(), p.ddep.as<int>(), ecnt.as<int>());
        HALA_CHECK_LAUNCH();
        exclusive_scan_i32(c, ecnt.as<int>(), epos.as<int>(), nslots);
        const int next = fetch_i32(c, epos.as<int>() + nslots);
        ci.reserve((size_t)ncell * CI * sizeof(int));
        colrange.reserve((size_t)2 * ncol * sizeof(int));
        init_range<<<(ncol + 63) / 64, 64, 0, c->stream>>>(ncol, nlev, colrange.as<int>());
        HALA_HIP(hipMemsetAsync(maxext.p, 0, sizeof(int), c->stream));
        col_cells<<<(int)((ncell + 255) / 256), 256, 0, c->stream>>>(nlev, ncol, p.lvl_chunk.as<int>(),
                                                                    cellcnt.as<int>(), cstart.as<int>(),
                                                                    epos.as<int>(), ci.as<int>(),
                                                                    colrange.as<int>(), maxext.as<int>());
        HALA_CHECK_LAUNCH();
        std::vector<int> hr(2 * ncol);
        HALA_HIP(hipMemcpyAsync(hr.data(), colrange.p, hr.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HALA_HIP(hipStreamSynchronize(c->stream));
        int nlevcol = 0;
        for (int k = 0; k < ncol; k++) nlevcol = std::max(nlevcol, hr[2 * k + 1] - hr[2 * k] + 1);
        p.col.why = 4;
        if (lds_bytes(K, CP, 2, nlevcol, sizeof(double)) > LDS_BUDGET) continue;
        // dependency records, level by level in k-major order
        tri_colplan &q = p.col;
        q.ncol = ncol;
        q.split = B;
        q.plane = B > 1 ? P : 0;
        q.K = K;
        q.CP = CP;
        q.nlevcol = nlevcol;
        q.L = L;
        q.nrec = (int64_t)nslots * q.L;
        q.next = next;
        q.code.reserve((size_t)q.nrec * sizeof(int));
        q.src.reserve((size_t)q.nrec * sizeof(int));
        fill_i32_k<<<(int)((q.nrec + 255) / 256), 256, 0, c->stream>>>(q.code.as<int>(), q.nrec, -1);
        fill_i32_k<<<(int)((q.nrec + 255) / 256), 256, 0, c->stream>>>(q.src.as<int>(), q.nrec, -1);
        HALA_CHECK_LAUNCH();
        HALA_CHECK_LAUNCH();
        q.spar.reserve((size_t)(nslots > 0 ? nslots : 1) * sizeof(int));
        fill_i32_k<<<(nslots + 255) / 256, 256, 0, c->stream>>>(q.spar.as<int>(), nslots, 0);
        q.ext_slot.reserve((size_t)(next + BTMAX) * sizeof(int));
        HALA_HIP(hipMemsetAsync(q.ext_slot.p, 0, q.ext_slot.bytes, c->stream));
        col_records<<<g, 256, 0, c->stream>>>(n, cmap.as<int>(), ncol, K, CP, q.L, nlev, p.lev.as<int>(),
                                              p.lvl_chunk.as<int>(), cstart.as<int>(), p.slot_of_row.as<int>(),
                                              p.dptr.as<int>(), p.ddep.as<int>(), p.dsrc.as<int>(), epos.as<int>(),
                                              colrange.as<int>(), q.code.as<int>(), q.src.as<int>(),
                                              q.ext_slot.as<int>(), q.spar.as<int>());
        HALA_CHECK_LAUNCH();
        // each external value staged once per cell
        {
            tmp_scope ts2(c);
            dbuf &uidx = ts2.get(), &fail = ts2.get();
            uidx.reserve((size_t)(next > 0 ? next : 1) * sizeof(int));
            fail.reserve(sizeof(int));
            HALA_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), c->stream));
            const int rawmax = fetch_i32(c, maxext.as<int>());  // the largest cell's raw external count
            int cap = 1;
            while (cap < rawmax && cap < DEDUP_MAX) cap <<= 1;
            HALA_HIP(hipMemsetAsync(maxext.p, 0, sizeof(int), c->stream));
            col_dedup<<<(unsigned)ncell, 256, (size_t)3 * cap * sizeof(int), c->stream>>>(
                nlev, ncol, cap, ci.as<int>(), q.ext_slot.as<int>(), uidx.as<int>(), maxext.as<int>(),
                fail.as<int>());
            HALA_CHECK_LAUNCH();
            col_fixcodes<<<(int)((q.nrec + 255) / 256), 256, 0, c->stream>>>(q.nrec, q.code.as<int>(),
                                                                             uidx.as<int>());
            HALA_CHECK_LAUNCH();
            const int mext = fetch_i32(c, maxext.as<int>());
            q.maxext = mext;
            // loader capacities built: 2 or 4 values per lane with the short-row
            // records, 4 or 8 with the long-row ones
            q.EL = ext_lanes(mext);
            if (q.L > 4 && q.EL == 2) q.EL = 4;
            if (q.L == 4 && q.EL > 4) q.EL = 0;
            const bool dfail = fetch_i32(c, fail.as<int>()) != 0;
            if (dfail || q.EL == 0) {
                tri_colplan_reset(q);
                q.why = dfail ? 5 : 6;
                continue;
            }
            if (lds_bytes(K, CP, q.EL, nlevcol, sizeof(double)) > LDS_BUDGET) {
                tri_colplan_reset(q);
                q.why = 4;
                continue;
            }
        }
        q.ci.swap(ci);
        q.colrange.swap(colrange);
        q.xs64.reserve((size_t)(nslots + BTMAX) * sizeof(double));
        HALA_HIP(hipStreamSynchronize(c->stream));
        q.ok = true;
        q.why = 0;
        return true;
    }
    return false;
}

void col_load_values(hala_ctx c, tri_plan &p, const double *factor, bool f32) {
    tri_colplan &q = p.col;
    if (!q.ok) return;
    const int nslots = p.nchunk * WAVE, npad = nslots + BTMAX;
    const int RB = rec_bytes(q.L, f32 ? sizeof(float) : sizeof(double));
    const int kcp = q.K * q.CP, xc = WAVE * q.EL;
    dbuf &dst = f32 ? q.val32 : q.val64;
    dst.reserve((size_t)npad * RB);
    if (!f32) q.hand64 = false;  // rewritten below: hand-off slots gone
    const int g = (int)(((int64_t)npad * (RB / 16) + 255) / 256);  // a thread per 16-byte quad
    if (f32) {
        col_pack<float><<<g, 256, 0, c->stream>>>(nslots, npad, q.L, RB, kcp, xc, q.code.as<int>(), q.spar.as<int>(),
                                                 q.src.as<int>(),
                                                 p.dg_src.as<int>(), factor, dst.as<unsigned char>());
        if (!q.xs32.p) q.xs32.reserve((size_t)npad * sizeof(float));
    } else {
        col_pack<double><<<g, 256, 0, c->stream>>>(nslots, npad, q.L, RB, kcp, xc, q.code.as<int>(), q.spar.as<int>(),
                                                  q.src.as<int>(),
                                                  p.dg_src.as<int>(), factor, dst.as<unsigned char>());
    }
    HALA_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------
namespace {

// Before a column sweep, over the whole chip: the sweep's polled output back
// to sentinels and the input into slot order.
// (Slots run level by level; a line of the input holds rows of ~16
// consecutive levels, so blocks are spread XCD-contiguously to reuse it in L2.)
template <class T, class TIN>
__global__ void col_gather_in(int nslots, const int *__restrict__ chunk_row, const TIN *__restrict__ in,
                              T *__restrict__ xs, T *__restrict__ out, const int *gate) {
    if (gate && *gate) return;
    const int s = xcd_block((nslots + 255) / 256) * 256 + threadIdx.x;
    if (s >= nslots) return;
    out[s] = sentinel<T>();
    const int r = chunk_row[s];
    if (r >= 0) xs[s] = (T)in[r];
}

// After a column sweep: the slot-ordered result into natural order.
template <class T, class TOUT>
__global__ void col_scatter_out(int n, const int *__restrict__ slot_of_row, const T *__restrict__ out,
                                TOUT *__restrict__ out_nat, const int *gate) {
    if (gate && *gate) return;
    const int r = xcd_block((n + 255) / 256) * 256 + threadIdx.x;
    if (r < n) out_nat[r] = (TOUT)out[slot_of_row[r]];
}

template <class T, class TOUT>
struct col_args {
    int nlev, K, CP;
    const int *ci, *colrange;
    const int4 *rec;   // packed row records (see col_pack), NQ quads per slot
    const int *ext_slot;
    const T *xs;  // input in slot order
    T alpha;
    T *out;
    T *hand;              // HAND: the next sweep's input, at the slot each record carries (col_set_handoff)
    int ncol;             // columns; workgroups past them are helpers
    T *arm;               // helpers: re-arm arm[0 .. narm) with sentinels (another sweep's polled output)
    int narm;
    col_axpy ax;          // helpers: ax.x += *ax.alpha * ax.p over ax.n rows (ax.x null: none)
    unsigned *ticket;  // [0] column ticket, [1] workgroups done
    int *err;
    const int *gate;
};

template <class T, int MAXD>
struct col_rec {
    static constexpr int NQ = (2 * MAXD + (int)sizeof(T) * (MAXD + 1) + 15) / 16;
    int slot;   // -1 for a lane without a row
    T xin;
    int4 q[NQ];  // packed record: int16 codes [MAXD] | diagonal | values [MAXD]
};

// 32-bit word w of a packed record (static w: a register pick)
template <int NQ>
__device__ __forceinline__ int rec_word(const int4 (&q)[NQ], int w) {
    const int4 v = q[w >> 2];
    switch (w & 3) {
    case 0: return v.x;
    case 1: return v.y;
    case 2: return v.z;
    default: return v.w;
    }
}
template <int NQ>
__device__ __forceinline__ int rec_code(const int4 (&q)[NQ], int k) {
    const int w = rec_word(q, k >> 1);
    return (k & 1) ? (w >> 16) : (int)(short)(w & 0xffff);
}
template <class T, int NQ>
__device__ __forceinline__ T rec_val(const int4 (&q)[NQ], int byte) {
    if (sizeof(T) == 8)
        return (T)__hiloint2double(rec_word(q, byte / 4 + 1), rec_word(q, byte / 4));
    return (T)__int_as_float(rec_word(q, byte / 4));
}

// Slow path, kept out of line so that the steady-state loop has no loops or
// conditional memory operations (the waits for the prefetched registers then
// stay precise): spin until the value is published, bounded.
template <class T>
__device__ __attribute__((noinline)) T poll_slow(const T *p, int *err) {
    T v = ld_agent(p);
    unsigned spins = 0;
    while (is_sent(v)) {
        __builtin_amdgcn_s_sleep(1);
        v = ld_agent(p);
        ++spins;
        if (spins > SPIN_LIMIT || ((spins & 1023u) == 0 && ld_agent(err) != 0)) {
            atomicOr(err, 1);
            return T(0);
        }
    }
    return v;
}

__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One workgroup per column: CP/64 compute waves (one row of a cell per
// thread) and one loader wave.  Relative level lam = level - first.
//
// Compute waves: rec[lam % P] is the row record of level lam and its input
// value, loaded P levels ahead.  These are their only global loads, so the
// wait for a record never waits for anything younger.  (The input is gathered
// into slot order by a pre-pass over the whole chip: read through the row map
// inside this kernel, each load instruction would touch 64 cache lines on the
// critical path.)
//
// Loader wave: lane l stages the cell's distinct externals l, l + 64, ...
// (EL per lane).  An external is written by another column only shortly
// before it is needed, so it is polled PE levels ahead (ev/evs rings), with
// its slot read from the external list P levels before that (esr ring).  The
// loader is a wave of its own because a wave's loads retire in issue order:
// a compute wave that also polled would wait, every level, for a poll issued
// one or two levels earlier, and the record prefetch would be lost.
//
// extb[lam & 1] holds the externals of level lam in LDS; a level ends with an
// LDS-only wait and the workgroup barrier.  Every step of either loop issues
// the same memory operations (idle lanes read valid padding), and the level
// count is padded to a multiple of P with empty levels, so both loop bodies
// are straight-line code.
//
// HAND (unit triangles only: the record's diagonal field holds the row's slot
// in the next sweep's input instead): each result is also stored into the
// next sweep's input, so that it needs no gather pass of its own.  (Rows of
// one cell and level lie close together in that order; a store in natural
// order would touch a line per lane on the critical path, so the natural-order
// result stays a pass of its own.)  Workgroups past the columns (their
// tickets come after every column's) are helpers: they re-arm another sweep's
// polled output on CUs the columns leave idle.
template <class T, class TOUT, int MODE, bool UNIT, bool HAND, int MAXD, int P, int PE, int EL, int BT>
__global__ __launch_bounds__(BT) void tri_col_kernel(col_args<T, TOUT> a) {
    static_assert(!HAND || UNIT, "the hand-off slot takes the diagonal's place");
    static_assert(P % PE == 0, "poll ring must divide the record ring");
    constexpr int XC = WAVE * EL;  // external values per level
    if (a.gate && *a.gate) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int CP = a.CP, K = a.K;
    T *win = reinterpret_cast<T *>(smem);
    T *extb = win + (size_t)K * CP;
    const int zidx = K * CP + 2 * XC;                         // win[zidx] = 0: the operand of no dependency
    int *eslot = reinterpret_cast<int *>(extb + 2 * XC + 2);  // slots of the polls in flight
    int *cis = eslot + PE * XC;
    __shared__ int col_s;
    const int t = threadIdx.x;
    if (t == 0) {
        col_s = (int)atomicAdd(a.ticket, 1u);
        win[zidx] = T(0);
    }
    __syncthreads();
    const int col = col_s;
    if (col >= a.ncol) {  // helper
        const int nh = (int)gridDim.x - a.ncol;
        for (int s = (col - a.ncol) * (int)blockDim.x + t; s < a.narm; s += nh * (int)blockDim.x)
            a.arm[s] = sentinel<T>();
        if (a.ax.x) {
            const double al = *a.ax.alpha;
            for (int i = (col - a.ncol) * (int)blockDim.x + t; i < a.ax.n; i += nh * (int)blockDim.x)
                a.ax.x[i] = a.ax.x[i] + al * a.ax.p[i];
        }
        if (t == 0) {
            unsigned d = atomicAdd(a.ticket + 1, 1u);
            if (d == gridDim.x - 1) {
                st_agent(a.ticket, 0u);
                st_agent(a.ticket + 1, 0u);
            }
        }
        return;
    }
    const int first = a.colrange[2 * col], last = a.colrange[2 * col + 1];

    const int nl = last - first + 1;
    const int nlp = (nl + P - 1) / P * P;  // padded level count
    const int ncis = nlp + 2 * P + 2;      // read-ahead reaches lam + PE + P + 1
    for (int i = t; i < ncis * CI; i += blockDim.x)
        cis[i] = (i < nl * CI) ? a.ci[((int64_t)col * a.nlev + first) * CI + i] : 0;
    __syncthreads();

    // wave index made scalar: each wave takes one branch as a whole
    const int wave = __builtin_amdgcn_readfirstlane(t / WAVE);
    if (wave == CP / WAVE) {
        // ---------------- loader wave ----------------
        const int l = t - CP;
        auto has = [&](int lam, int e) { return l + WAVE * e < (cis[lam * CI + 2] >> 16); };
        // slot of external l + 64 e of level lam (a valid slot even past the cell's list)
        auto raw = [&](int lam, int e) { return a.ext_slot[cis[lam * CI + 3] + l + WAVE * e]; };
        // polled address: workgroup-uniform base + 32-bit byte offset (no 64-bit
        // register pair per slot, so a freshly loaded slot is used where it lands)
        const char *obase = reinterpret_cast<const char *>(a.out);
        auto at = [&](int slot) { return reinterpret_cast<const T *>(obase + (unsigned)slot * (unsigned)sizeof(T)); };
        T ev[PE][EL];
        int esr[P][EL];
        // level 0 directly; polls of levels 1..PE; slots of levels PE+1 .. PE+P
#pragma unroll
        for (int e = 0; e < EL; e++)
            if (has(0, e)) extb[l + WAVE * e] = poll_slow(at(raw(0, e)), a.err);
#pragma unroll
        for (int i = 0; i < PE; i++)
#pragma unroll
            for (int e = 0; e < EL; e++) {
                const int sl = raw(i + 1, e);
                eslot[i * XC + l + WAVE * e] = sl;
                ev[i][e] = ld_agent(at(sl));
            }
#pragma unroll
        for (int i = 0; i < P; i++)
#pragma unroll
            for (int e = 0; e < EL; e++) esr[i][e] = raw(PE + 1 + i, e);
        lds_barrier();
        for (int lam0 = 0; lam0 < nlp; lam0 += P) {
#pragma unroll
            for (int s = 0; s < P; s++) {
                const int lam = lam0 + s;
                const int se = s % PE;  // static once unrolled
                // externals of level lam + 1 into LDS
                T *dst = extb + ((lam + 1) & 1) * XC + l;
                bool slow = false;
#pragma unroll
                for (int e = 0; e < EL; e++) {
                    T v = ev[se][e];
                    if (has(lam + 1, e) && is_sent(v)) {
                        v = poll_slow(at(eslot[se * XC + l + WAVE * e]), a.err);
                        slow = true;
                    }
                    dst[WAVE * e] = v;
                }
                // A miss means this column has caught up with a producer: every poll
                // already in flight was issued too early as well.  Waiting for the next
                // level's values too puts the column PE levels behind again, so the
                // polls issued from here on find their values written.
                if (PE > 1 && __ballot(slow) != 0) {
                    const int sn = (se + 1) % PE;
#pragma unroll
                    for (int e = 0; e < EL; e++) {
                        T v = ev[sn][e];
                        if (has(lam + 2, e) && is_sent(v)) v = poll_slow(at(eslot[sn * XC + l + WAVE * e]), a.err);
                        ev[sn][e] = v;
                    }
                }
                // polls of level lam + PE + 1, slots of level lam + PE + P + 1
#pragma unroll
                for (int e = 0; e < EL; e++) {
                    eslot[se * XC + l + WAVE * e] = esr[s][e];
                    ev[se][e] = ld_agent(at(esr[s][e]));
                }
#pragma unroll
                for (int e = 0; e < EL; e++) esr[s][e] = raw(lam + PE + P + 1, e);
                lds_barrier();
                // keep the next level's first look at its polled values below the
                // barrier: hoisted above it, the loader would wait there for a poll
                // issued one level earlier instead of two
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    } else {
        // ---------------- compute waves ----------------
        constexpr int NQ = col_rec<T, MAXD>::NQ;
        // 32-bit byte offsets from uniform bases (the plan keeps them below 2^31)
        auto load_rec_at = [&](col_rec<T, MAXD> &r, int slot0, int size) {
            const unsigned s = (unsigned)(slot0 + t);
            r.slot = t < size ? (int)s : -1;
            // slot arrays are padded past their end: idle lanes read in bounds
            r.xin = *reinterpret_cast<const T *>(reinterpret_cast<const char *>(a.xs) + s * (unsigned)sizeof(T));
            const int4 *src = reinterpret_cast<const int4 *>(reinterpret_cast<const char *>(a.rec) + s * (unsigned)(16 * NQ));
#pragma unroll
            for (int i = 0; i < NQ; i++) r.q[i] = src[i];
        };
        auto load_rec = [&](col_rec<T, MAXD> &r, int lam) {
            const int *ci = cis + lam * CI;
            load_rec_at(r, ci[0], ci[2] & 0xffff);
        };
        col_rec<T, MAXD> rec[P];
        // in the loop's own issue order, so that the loop is entered with the
        // outstanding loads it keeps in steady state
#pragma unroll
        for (int i = 0; i < P; i++) {
            load_rec(rec[i], i);
            __builtin_amdgcn_sched_barrier(0);
        }
        lds_barrier();
        for (int lam0 = 0; lam0 < nlp; lam0 += P) {
#pragma unroll
            for (int s = 0; s < P; s++) {
                const int lam = lam0 + s;
                const int lvl = first + lam;
                const col_rec<T, MAXD> &r = rec[s];
                const bool valid = r.slot >= 0;
                T acc = (MODE == TRI_SUB) ? r.xin : T(0), sum = T(0);
                // all LDS reads of the step (the cell record of level lam + P and
                // the row's operands) issued before the first is waited for
                const int *cin = cis + (lam + P) * CI;
                const int nslot0 = cin[0], nsize = cin[2] & 0xffff;
                int cds[MAXD];
                T xs[MAXD];
#pragma unroll
                for (int k = 0; k < MAXD; k++) {
                    const int cd = rec_code(r.q, k);  // an LDS element (see col_pack)
                    cds[k] = cd;
                    xs[k] = win[cd];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int k = 0; k < MAXD; k++) {
                    const int cd = cds[k];
                    const T v = rec_val<T>(r.q, 2 * MAXD + (int)sizeof(T) * (k + 1));
                    const T x = xs[k];
                    if (MODE == TRI_SUB) {
                        // no dependency: acc - 0 * 0, which is acc exactly (also for -0)
                        acc = acc - v * x;
                    } else {
                        // a sum starting at +0 would turn -0 into +0: skip instead
                        const T nsum = sum + v * x;
                        sum = cd != zidx ? nsum : sum;
                    }
                }
                if (MODE == TRI_SUM) acc = a.alpha * r.xin - sum;
                if (!UNIT) acc = acc / rec_val<T>(r.q, 2 * MAXD);
                win[(lvl % K) * CP + t] = acc;
                if (valid)
                    st_agent(reinterpret_cast<T *>(reinterpret_cast<char *>(a.out) + (unsigned)r.slot * (unsigned)sizeof(T)),
                             acc);
                if (HAND && valid)
                    *reinterpret_cast<T *>(reinterpret_cast<char *>(a.hand) +
                                           (unsigned)rec_word(r.q, MAXD / 2) * (unsigned)sizeof(T)) = acc;
                load_rec_at(rec[s], nslot0, nsize);
                lds_barrier();
            }
        }
        if (t == 0) {
            unsigned d = atomicAdd(a.ticket + 1, 1u);
            if (d == gridDim.x - 1) {
                st_agent(a.ticket, 0u);
                st_agent(a.ticket + 1, 0u);
            }
        }
    }
}

constexpr int ARM_HELPERS = 64;  // helper workgroups of a sweep with side work (re-arming, vector update)

template <class T, class TOUT, int MODE, bool UNIT, bool HAND, int MAXD, int P, int EL>
void launch_col_el(hala_ctx c, const tri_colplan &q, col_args<T, TOUT> &a, size_t lds) {
    constexpr int PE = poll_ahead(MAXD);
    static_assert(PE <= PE_MAX, "LDS sized for PE_MAX polls in flight");
    constexpr int BT = (MAXD == 4 ? CPMAX : CPMAX_LONG) + WAVE;
    const int nwg = q.ncol + (a.narm > 0 || a.ax.x ? ARM_HELPERS : 0);
    tri_col_kernel<T, TOUT, MODE, UNIT, HAND, MAXD, P, PE, EL, BT><<<dim3(nwg), dim3(q.CP + WAVE), lds, c->stream>>>(a);
}

template <class T, class TOUT, int MODE, bool UNIT, bool HAND = false>
void launch_col(hala_ctx c, tri_plan &p, col_args<T, TOUT> &a, size_t lds) {
    const tri_colplan &q = p.col;
    if (q.L == 4) {
        if (q.EL == 2) launch_col_el<T, TOUT, MODE, UNIT, HAND, 4, 8, 2>(c, q, a, lds);
        else launch_col_el<T, TOUT, MODE, UNIT, HAND, 4, 8, 4>(c, q, a, lds);
    } else {
        if (q.EL == 4) launch_col_el<T, TOUT, MODE, UNIT, HAND, 16, 4, 4>(c, q, a, lds);
        else launch_col_el<T, TOUT, MODE, UNIT, HAND, 16, 4, 8>(c, q, a, lds);
    }
}

}  // namespace

bool col_usable(hala_ctx c, const tri_plan &p, bool is64) {
    const tri_colplan &q = p.col;
    if (!q.ok || !c->opt.tri_col) return false;
    if (q.CP > (q.L == 4 ? CPMAX : CPMAX_LONG)) return false;
    // the kernel addresses records, inputs and outputs by 32-bit byte offsets
    const int64_t npad = (int64_t)p.nchunk * WAVE + BTMAX;
    if (npad * rec_bytes(q.L, sizeof(double)) > INT_MAX) return false;
    return is64 ? (q.val64.p != nullptr) : (q.val32.p != nullptr);
}

namespace {

// The column kernel alone: input already in slot order in xs; the result in
// slot order in the polled output and, with next (FP64, TRI_SUB, unit p with
// hand-off slots), in next's slot-order input; arm (nullable, narm slots) is
// re-armed by helper workgroups.
template <class T, class TOUT>
void col_sweep(hala_ctx c, tri_plan &p, int mode, T alpha, const int *gate, tri_plan *next, T *arm, int narm,
               const col_axpy *ax = nullptr) {
    constexpr bool is64 = sizeof(T) == 8;
    tri_colplan &q = p.col;
    col_args<T, TOUT> a;
    a.nlev = p.nlev;
    a.K = q.K;
    a.CP = q.CP;
    a.ci = q.ci.as<int>();
    a.colrange = q.colrange.as<int>();
    a.rec = is64 ? reinterpret_cast<const int4 *>(q.val64.p) : reinterpret_cast<const int4 *>(q.val32.p);
    a.ext_slot = q.ext_slot.as<int>();
    a.xs = is64 ? reinterpret_cast<T *>(q.xs64.p) : reinterpret_cast<T *>(q.xs32.p);
    a.alpha = alpha;
    a.out = is64 ? reinterpret_cast<T *>(p.out64.p) : reinterpret_cast<T *>(p.out32.p);
    a.hand = nullptr;
    a.ncol = q.ncol;
    a.arm = arm;
    a.narm = arm ? narm : 0;
    a.ax = ax ? *ax : col_axpy{nullptr, nullptr, nullptr, 0};
    a.ticket = counters(c) + 104;
    a.err = wait_error_flag(c);
    a.gate = gate;
    const size_t lds = (size_t)lds_bytes(q.K, q.CP, q.EL, q.nlevcol, sizeof(T));
    if (next) {
        // the records carry next's input slots (FP64 records of a unit triangle)
        HALA_REQUIRE(is64 && mode == TRI_SUB && p.unit && q.hand64, HALA_ERR_ARG, "column sweep: no hand-off slots");
        a.hand = reinterpret_cast<T *>(next->col.xs64.p);
        if constexpr (is64) launch_col<T, TOUT, TRI_SUB, true, true>(c, p, a, lds);
    } else if (mode == TRI_SUB) {
        if (p.unit) launch_col<T, TOUT, TRI_SUB, true>(c, p, a, lds);
        else launch_col<T, TOUT, TRI_SUB, false>(c, p, a, lds);
    } else {
        if (p.unit) launch_col<T, TOUT, TRI_SUM, true>(c, p, a, lds);
        else launch_col<T, TOUT, TRI_SUM, false>(c, p, a, lds);
    }
    HALA_CHECK_LAUNCH();
}

}  // namespace

template <class T, class TOUT>
void tri_col_exec(hala_ctx c, tri_plan &p, int mode, T alpha, const TOUT *in, TOUT *out_nat, const int *gate,
                  tri_plan *next) {
    constexpr bool is64 = sizeof(T) == 8;
    tri_colplan &q = p.col;
    const int nslots = p.nchunk * WAVE;
    T *xs = is64 ? reinterpret_cast<T *>(q.xs64.p) : reinterpret_cast<T *>(q.xs32.p);
    T *out = is64 ? reinterpret_cast<T *>(p.out64.p) : reinterpret_cast<T *>(p.out32.p);
    col_gather_in<T, TOUT><<<xcd_grid((nslots + 255) / 256), 256, 0, c->stream>>>(nslots, p.chunk_row.as<int>(), in,
                                                                                 xs, out, gate);
    HALA_CHECK_LAUNCH();
    T *arm = nullptr;
    int narm = 0;
    if (next) {  // next's polled output re-armed by this sweep's helpers
        arm = is64 ? reinterpret_cast<T *>(next->out64.p) : reinterpret_cast<T *>(next->out32.p);
        narm = next->nchunk * WAVE;
    }
    col_sweep<T, TOUT>(c, p, mode, alpha, gate, next, arm, narm);
    if (out_nat) {
        col_scatter_out<T, TOUT><<<xcd_grid((p.n + 255) / 256), 256, 0, c->stream>>>(p.n, p.slot_of_row.as<int>(), out,
                                                                                    out_nat, gate);
        HALA_CHECK_LAUNCH();
    }
}

void col_sweep64(hala_ctx c, tri_plan &p, int mode, double alpha, const int *gate, tri_plan *next, tri_plan *arm,
                 const col_axpy *ax) {
    col_sweep<double, double>(c, p, mode, alpha, gate, next, arm ? arm->out64.as<double>() : nullptr,
                              arm ? arm->nchunk * WAVE : 0, ax);
}

void col_scatter64(hala_ctx c, tri_plan &p, double *out_nat, const int *gate) {
    col_scatter_out<double, double><<<xcd_grid((p.n + 255) / 256), 256, 0, c->stream>>>(
        p.n, p.slot_of_row.as<int>(), p.out64.as<double>(), out_nat, gate);
    HALA_CHECK_LAUNCH();
}

// The slot of every row of `src` in the input of `dst` (a sink slot past
// dst's slots for record lanes without a row), written into the diagonal field
// of src's FP64 records, which a unit triangle does not read.
__global__ void col_set_hand(int nslots, const int *__restrict__ chunk_row, const int *__restrict__ dst_slot,
                             int sink, int RB, int word, unsigned char *__restrict__ rec) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int r = chunk_row[s];
    int *w = reinterpret_cast<int *>(rec + (size_t)s * RB) + word;
    w[0] = r >= 0 ? dst_slot[r] : sink;
    w[1] = 0;
}

void col_set_handoff(hala_ctx c, tri_plan &src, tri_plan &dst) {
    tri_colplan &q = src.col;
    q.hand64 = false;
    if (!q.ok || !dst.col.ok || !src.unit || !q.val64.p || !dst.col.xs64.p) return;
    const int nslots = src.nchunk * WAVE;
    col_set_hand<<<(nslots + 255) / 256, 256, 0, c->stream>>>(nslots, src.chunk_row.as<int>(),
                                                             dst.slot_of_row.as<int>(), dst.nchunk * WAVE,
                                                             rec_bytes(q.L, sizeof(double)), q.L / 2,
                                                             q.val64.as<unsigned char>());
    HALA_CHECK_LAUNCH();
    q.hand64 = true;
}

template void tri_col_exec<double, double>(hala_ctx, tri_plan &, int, double, const double *, double *, const int *,
                                           tri_plan *);
template void tri_col_exec<float, double>(hala_ctx, tri_plan &, int, float, const double *, double *, const int *,
                                          tri_plan *);
template void tri_col_exec<float, float>(hala_ctx, tri_plan &, int, float, const float *, float *, const int *,
                                         tri_plan *);

}  // namespace hala_amd
