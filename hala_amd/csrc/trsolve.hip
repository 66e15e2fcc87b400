// trsolve.hip -- plan construction and sync-free level-ordered execution.
#include <algorithm>
#include <climits>

#include "trsolve.hpp"

namespace hala_amd {


// ---------------------------------------------------------------------------
// dependency lists
// ---------------------------------------------------------------------------
__global__ void dep_count(int n, int sel, const int *__restrict__ pntr, const int *__restrict__ diagpos,
                          const int *__restrict__ tpntr, const int *__restrict__ trow, int *__restrict__ cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = 0;
    switch (sel) {
    case DEP_LOWER: c = diagpos[i] - pntr[i]; break;
    case DEP_UPPER: c = pntr[i + 1] - diagpos[i] - 1; break;
    case DEP_ROW_BUT_LAST:
    case DEP_ROW_BUT_FIRST: c = max(pntr[i + 1] - pntr[i] - 1, 0); break;
    default:
        for (int e = tpntr[i]; e < tpntr[i + 1]; e++) {
            int r = trow[e];
            if (sel == DEP_T_LOWER ? (r < i) : (r > i)) c++;
        }
    }
    cnt[i] = c;
}

__global__ void dep_fill(int n, int sel, const int *__restrict__ pntr, const int *__restrict__ indx,
                         const int *__restrict__ diagpos, const int *__restrict__ tpntr,
                         const int *__restrict__ trow, const int *__restrict__ tmap, const int *__restrict__ dptr,
                         int *__restrict__ ddep, int *__restrict__ dsrc, int *__restrict__ dgs, int unit) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int q = dptr[i];
    switch (sel) {
    case DEP_LOWER:
        for (int j = pntr[i]; j < diagpos[i]; j++, q++) ddep[q] = indx[j], dsrc[q] = j;
        break;
    case DEP_UPPER:
        for (int j = diagpos[i] + 1; j < pntr[i + 1]; j++, q++) ddep[q] = indx[j], dsrc[q] = j;
        break;
    case DEP_ROW_BUT_LAST:
        for (int j = pntr[i]; j < pntr[i + 1] - 1; j++, q++) ddep[q] = indx[j], dsrc[q] = j;
        break;
    case DEP_ROW_BUT_FIRST:
        for (int j = pntr[i] + 1; j < pntr[i + 1]; j++, q++) ddep[q] = indx[j], dsrc[q] = j;
        break;
    case DEP_T_LOWER:
    case DEP_T_UPPER_ASC:
        for (int e = tpntr[i]; e < tpntr[i + 1]; e++) {
            int r = trow[e];
            if (sel == DEP_T_LOWER ? (r < i) : (r > i)) ddep[q] = r, dsrc[q] = tmap[e], q++;
        }
        break;
    case DEP_T_UPPER:
        for (int e = tpntr[i + 1] - 1; e >= tpntr[i]; e--) {
            int r = trow[e];
            if (r > i) ddep[q] = r, dsrc[q] = tmap[e], q++;
        }
        break;
    }
    int d = -1;
    if (!unit) {
        if (sel == DEP_ROW_BUT_LAST) d = pntr[i + 1] - 1;
        else if (sel == DEP_ROW_BUT_FIRST) d = pntr[i];
        else d = diagpos[i];
    }
    dgs[i] = d;
}

void build_deps(hala_ctx c, tri_plan &p, int n, dep_sel sel, const int *pntr, const int *indx, const int *diagpos,
                const csr_transpose *tp, bool unit, bool backward) {
    p.n = n;
    p.unit = unit;
    p.backward = backward;
    const int *tpntr = tp ? tp->tpntr.as<int>() : nullptr;
    const int *trow = tp ? tp->trow.as<int>() : nullptr;
    const int *tmap = tp ? tp->tmap.as<int>() : nullptr;
    tmp_scope ts(c);
    dbuf &cnt = ts.get();
    cnt.reserve((size_t)(n + 1) * sizeof(int));
    p.dptr.reserve((size_t)(n + 1) * sizeof(int));
    int g = (n + 255) / 256;
    dep_count<<<g, 256, 0, c->stream>>>(n, sel, pntr, diagpos, tpntr, trow, cnt.as<int>());
    HALA_CHECK_LAUNCH();
    exclusive_scan_i32(c, cnt.as<int>(), p.dptr.as<int>(), n);
    int total = fetch_i32(c, p.dptr.as<int>() + n);
    p.ddep.reserve((size_t)(total > 0 ? total : 1) * sizeof(int));
    p.dsrc.reserve((size_t)(total > 0 ? total : 1) * sizeof(int));
    p.diag_src.reserve((size_t)n * sizeof(int));
    dep_fill<<<g, 256, 0, c->stream>>>(n, sel, pntr, indx, diagpos, tpntr, trow, tmap, p.dptr.as<int>(),
                                       p.ddep.as<int>(), p.dsrc.as<int>(), p.diag_src.as<int>(), unit ? 1 : 0);
    HALA_CHECK_LAUNCH();
    HALA_HIP(hipStreamSynchronize(c->stream));
}

// ---------------------------------------------------------------------------
// level sets: sync-free, rows in order (forward) or reverse order (backward)
// ---------------------------------------------------------------------------
// A lane keeps LDEP dependencies of its row in registers and polls all of
// those still pending at once, so a row waits one memory round trip per
// attempt rather than one per dependency.
// Wavefront w takes rows w*64 .. w*64+63 (in sweep order) by its index:
// a row depends only on rows of lower wavefronts (or its own), and
// workgroups are dispatched in index order on each XCD, so the lowest
// unfinished wavefront has what it waits for stored, or in its own lanes.
// (An atomic ticket per wavefront would not need the dispatch order, but
// queues every wavefront's start on one counter.)
constexpr int LDEP = 8;
__global__ __launch_bounds__(256) void level_kernel(int n, int backward, const int *__restrict__ dptr,
                                                    const int *__restrict__ ddep, int *lev, int *err) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    int idx = w * WAVE + lane;
    int row = backward ? n - 1 - idx : idx;
    bool done = !(idx < n);
    int cb = 0, end = 0, m = 0;
    unsigned pend = 0;  // bit q: dependency cb + q not yet seen
    int dk[LDEP];
    auto load_chunk = [&] {
        pend = 0;
#pragma unroll
        for (int q = 0; q < LDEP; q++) {
            const bool in = cb + q < end;
            dk[q] = in ? ddep[cb + q] : 0;
            pend |= in ? (1u << q) : 0u;
        }
    };
    if (!done) {
        cb = dptr[row];
        end = dptr[row + 1];
        load_chunk();
    }
    unsigned spins = 0;
    while (__any(!done)) {
        if (!done) {
            int lv[LDEP];
#pragma unroll
            for (int q = 0; q < LDEP; q++) lv[q] = (pend >> q & 1u) ? ld_agent(lev + dk[q]) : -1;
#pragma unroll
            for (int q = 0; q < LDEP; q++)
                if ((pend >> q & 1u) && lv[q] >= 0) {
                    m = max(m, lv[q] + 1);
                    pend &= ~(1u << q);
                }
            if (pend == 0) {
                cb += LDEP;
                if (cb < end) {
                    load_chunk();
                } else {
                    st_agent(lev + row, m);
                    done = true;
                }
            }
        }
        if (__any(!done)) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > SPIN_LIMIT) {
                if (!done) {
                    atomicOr(err, 1);
                    st_agent(lev + row, 0);
                    done = true;
                }
            }
        }
    }
}

__global__ void fill_i32_k(int *__restrict__ x, int64_t n, int v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// Candidate levels from another triangle's (top = its level count - 1):
// lev[r] = top - lo[r], checked against this triangle's own level recurrence
// (0 for a row without dependencies, else 1 + the largest level among them).
// A row that breaks it sets *bad.  When none does, the candidate is the
// recurrence's unique solution, which level_kernel would compute.
__global__ void level_mirror(int n, int top, const int *__restrict__ lo, const int *__restrict__ dptr,
                             const int *__restrict__ ddep, int *__restrict__ lev, int *__restrict__ bad) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    int cand = 0, m = 0;
    if (r < n) {
        cand = top - lo[r];
        for (int e = dptr[r]; e < dptr[r + 1]; e++) m = max(m, top - lo[ddep[e]] + 1);
        lev[r] = cand;
    }
    // one store per wavefront, none once the flag is seen set
    if (__any(m != cand) && (threadIdx.x & (WAVE - 1)) == 0 && ld_agent(bad) == 0) *bad = 1;
}

template <class T>
__global__ void fill_sent_k(T *__restrict__ x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = sentinel<T>();
}

// chunk counts per level
__global__ void level_chunks(int nlev, const int *__restrict__ level_ptr, int *__restrict__ nch) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < nlev) nch[l] = (level_ptr[l + 1] - level_ptr[l] + WAVE - 1) / WAVE;
}

// position p in perm -> chunk slot
__global__ void place_rows(int n, const int *__restrict__ perm, const int *__restrict__ lev,
                           const int *__restrict__ level_ptr, const int *__restrict__ chunk_ptr,
                           int *__restrict__ chunk_row, int *__restrict__ slot_of_row) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int row = perm[p];
    int l = lev[row];
    int q = p - level_ptr[l];
    int slot = (chunk_ptr[l] + q / WAVE) * WAVE + (q % WAVE);
    chunk_row[slot] = row;
    slot_of_row[row] = slot;
}

// Longest dependency list of every chunk.  A chunk's 64 slots are exactly one
// wavefront of this launch (blocks of 256 slots), so the maximum is a lane
// reduction and one store, not 64 atomics on one counter.
__global__ __launch_bounds__(256) void chunk_lengths(int nslots, const int *__restrict__ chunk_row,
                                                     const int *__restrict__ dptr, int *__restrict__ chunk_len) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    int len = 0;
    if (s < nslots) {
        const int row = chunk_row[s];
        if (row >= 0) len = dptr[row + 1] - dptr[row];
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0 && s < nslots) chunk_len[s / WAVE] = len;
}

__global__ void fill_entries(int nslots, const int *__restrict__ chunk_row, const int *__restrict__ chunk_off,
                             const int *__restrict__ chunk_len, const int *__restrict__ dptr,
                             const int *__restrict__ ddep, const int *__restrict__ dsrc,
                             const int *__restrict__ slot_of_row, const int *__restrict__ diag_src,
                             int *__restrict__ ent_dep, int *__restrict__ ent_src, int *__restrict__ dg_src) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    int c = s / WAVE, lane = s % WAVE;
    int row = chunk_row[s];
    int len = chunk_len[c];
    int64_t base = (int64_t)chunk_off[c] * WAVE + lane;
    int k = 0;
    if (row >= 0) {
        int a = dptr[row], b = dptr[row + 1];
        for (; k < b - a; k++) {
            ent_dep[base + (int64_t)k * WAVE] = slot_of_row[ddep[a + k]];
            ent_src[base + (int64_t)k * WAVE] = dsrc[a + k];
        }
        dg_src[s] = diag_src[row];
    } else {
        dg_src[s] = -1;
    }
    for (; k < len; k++) {
        ent_dep[base + (int64_t)k * WAVE] = -1;
        ent_src[base + (int64_t)k * WAVE] = -1;
    }
}

namespace {

// Slot layout from a level-major row order: chunk rows and slot of every row,
// entry slabs (k-th dependency of the 64 lanes contiguous) and diagonal sources.
void lay_out_slots(hala_ctx c, tri_plan &p, const int *order) {
    const int n = p.n, nslots = p.nchunk * WAVE, g = (n + 255) / 256;
    p.chunk_row.reserve((size_t)(nslots + SLOT_PAD) * sizeof(int));
    p.slot_of_row.reserve((size_t)n * sizeof(int));
    p.chunk_len.reserve((size_t)(p.nchunk + 1) * sizeof(int));
    p.chunk_off.reserve((size_t)(p.nchunk + 1) * sizeof(int));
    fill_i32_k<<<(nslots + SLOT_PAD + 255) / 256, 256, 0, c->stream>>>(p.chunk_row.as<int>(), nslots + SLOT_PAD, -1);
    fill_i32_k<<<(p.nchunk + 256) / 256, 256, 0, c->stream>>>(p.chunk_len.as<int>(), p.nchunk + 1, 0);
    HALA_CHECK_LAUNCH();
    place_rows<<<g, 256, 0, c->stream>>>(n, order, p.lev.as<int>(), p.level_ptr.as<int>(), p.lvl_chunk.as<int>(),
                                         p.chunk_row.as<int>(), p.slot_of_row.as<int>());
    HALA_CHECK_LAUNCH();
    chunk_lengths<<<(nslots + 255) / 256, 256, 0, c->stream>>>(nslots, p.chunk_row.as<int>(), p.dptr.as<int>(),
                                                              p.chunk_len.as<int>());
    HALA_CHECK_LAUNCH();
    exclusive_scan_i32(c, p.chunk_len.as<int>(), p.chunk_off.as<int>(), p.nchunk);
    int slabs = fetch_i32(c, p.chunk_off.as<int>() + p.nchunk);
    p.nent = (int64_t)slabs * WAVE;
    size_t ne = (size_t)(p.nent > 0 ? p.nent : 1);
    p.ent_dep.reserve(ne * sizeof(int));
    p.ent_src.reserve(ne * sizeof(int));
    p.dg_src.reserve((size_t)nslots * sizeof(int));
    fill_entries<<<(nslots + 255) / 256, 256, 0, c->stream>>>(
        nslots, p.chunk_row.as<int>(), p.chunk_off.as<int>(), p.chunk_len.as<int>(), p.dptr.as<int>(),
        p.ddep.as<int>(), p.dsrc.as<int>(), p.slot_of_row.as<int>(), p.diag_src.as<int>(), p.ent_dep.as<int>(),
        p.ent_src.as<int>(), p.dg_src.as<int>());
    HALA_CHECK_LAUNCH();
    p.laid_out = true;
}

__global__ void level_key(int n, int nkey, const int *__restrict__ lev, const int *__restrict__ key,
                          int *__restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = lev[r] * nkey + key[r];
}

}  // namespace

void sort_level_keys(hala_ctx c, const tri_plan &p, const int *key, int nkey, int *bptr, int *order) {
    const int n = p.n;
    HALA_REQUIRE((int64_t)p.nlev * nkey < INT_MAX, HALA_ERR_ARG, "triangular plan: too many level keys");
    tmp_scope ts(c);
    dbuf &bk = ts.get();
    bk.reserve((size_t)n * sizeof(int));
    level_key<<<(n + 255) / 256, 256, 0, c->stream>>>(n, nkey, p.lev.as<int>(), key, bk.as<int>());
    HALA_CHECK_LAUNCH();
    bucket_sort(c, n, bk.as<int>(), p.nlev * nkey, bptr, order);
}

void apply_level_order(hala_ctx c, tri_plan &p, const int *order) {
    // the new order is also the plan's level order (rows of a bucket stay together)
    HALA_HIP(hipMemcpyAsync(p.perm.p, order, (size_t)p.n * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    p.perm_sorted = false;
    lay_out_slots(c, p, p.perm.as<int>());
    HALA_HIP(hipStreamSynchronize(c->stream));
}

void compute_levels(hala_ctx c, tri_plan &p) {
    const int n = p.n;
    HALA_REQUIRE(n > 0, HALA_ERR_ARG, "triangular plan: empty matrix");
    int g = (n + 255) / 256;
    p.lev.reserve((size_t)n * sizeof(int));
    fill_i32_k<<<g, 256, 0, c->stream>>>(p.lev.as<int>(), n, -1);
    HALA_CHECK_LAUNCH();
    int nwaves = (n + WAVE - 1) / WAVE;
    int lblocks = (nwaves + 3) / 4;
    level_kernel<<<lblocks, 256, 0, c->stream>>>(n, p.backward ? 1 : 0, p.dptr.as<int>(), p.ddep.as<int>(),
                                                 p.lev.as<int>(), wait_error_flag(c));
    HALA_CHECK_LAUNCH();
    check_wait_error(c);
    p.nlev = reduce_max_i32(c, p.lev.as<int>(), n) + 1;
}

bool mirror_levels(hala_ctx c, tri_plan &p, const int *lev_other, int nlev_other) {
    const int n = p.n;
    HALA_REQUIRE(n > 0 && nlev_other > 0, HALA_ERR_ARG, "triangular plan: empty matrix");
    tmp_scope ts(c);
    dbuf &bad = ts.get();
    bad.reserve(sizeof(int));
    HALA_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), c->stream));
    p.lev.reserve((size_t)n * sizeof(int));
    level_mirror<<<(n + 255) / 256, 256, 0, c->stream>>>(n, nlev_other - 1, lev_other, p.dptr.as<int>(),
                                                         p.ddep.as<int>(), p.lev.as<int>(), bad.as<int>());
    HALA_CHECK_LAUNCH();
    if (fetch_i32(c, bad.as<int>()) != 0) return false;
    p.nlev = nlev_other;  // the candidate's largest value: the other triangle has rows at level 0
    return true;
}

void build_schedule(hala_ctx c, tri_plan &p, bool want32, bool have_levels) {
    const int n = p.n;
    HALA_REQUIRE(n > 0, HALA_ERR_ARG, "triangular plan: empty matrix");
    if (!have_levels) compute_levels(c, p);
    // rows grouped by level, sorted by row inside a level (the column layout
    // may reorder every level its own way later)
    p.level_ptr.reserve((size_t)(p.nlev + 1) * sizeof(int));
    p.perm.reserve((size_t)n * sizeof(int));
    p.perm_sorted = true;
    p.laid_out = false;
    bucket_sort(c, n, p.lev.as<int>(), p.nlev, p.level_ptr.as<int>(), p.perm.as<int>());
    p.h_level_ptr.resize(p.nlev + 1);
    HALA_HIP(hipMemcpyAsync(p.h_level_ptr.data(), p.level_ptr.p, (size_t)(p.nlev + 1) * sizeof(int),
                            hipMemcpyDeviceToHost, c->stream));
    // chunks: ceil(level size / 64) per level
    tmp_scope ts(c);
    dbuf &nch = ts.get();
    dbuf &chunk_ptr = p.lvl_chunk;
    nch.reserve((size_t)(p.nlev + 1) * sizeof(int));
    chunk_ptr.reserve((size_t)(p.nlev + 1) * sizeof(int));
    level_chunks<<<(p.nlev + 255) / 256, 256, 0, c->stream>>>(p.nlev, p.level_ptr.as<int>(), nch.as<int>());
    HALA_CHECK_LAUNCH();
    exclusive_scan_i32(c, nch.as<int>(), chunk_ptr.as<int>(), p.nlev);
    p.nchunk = fetch_i32(c, chunk_ptr.as<int>() + p.nlev);
    const int nslots = p.nchunk * WAVE;
    // the column layout lays the slots out in its own order; else (or when it
    // gave up before) the level order does
    if (c->opt.tri_col) build_col_plan(c, p, c->opt.tri_col_cols);
    if (!p.laid_out) lay_out_slots(c, p, p.perm.as<int>());
    p.out64.reserve((size_t)nslots * sizeof(double));
    if (want32) p.out32.reserve((size_t)nslots * sizeof(float));
    tri_rearm(c, p);
    HALA_HIP(hipStreamSynchronize(c->stream));
}

void sort_perm(hala_ctx c, tri_plan &p) {
    if (p.perm_sorted) return;
    bucket_sort(c, p.n, p.lev.as<int>(), p.nlev, p.level_ptr.as<int>(), p.perm.as<int>());
    p.perm_sorted = true;
}

void tri_rearm_prec(hala_ctx c, tri_plan &p, bool is64) {
    int64_t nslots = (int64_t)p.nchunk * WAVE;
    int g = (int)((nslots + 255) / 256);
    if (is64) fill_sent_k<double><<<g, 256, 0, c->stream>>>(p.out64.as<double>(), nslots);
    else fill_sent_k<float><<<g, 256, 0, c->stream>>>(p.out32.as<float>(), nslots);
    HALA_CHECK_LAUNCH();
}

void tri_rearm(hala_ctx c, tri_plan &p) {
    int64_t nslots = (int64_t)p.nchunk * WAVE;
    int g = (int)((nslots + 255) / 256);
    if (p.out64.p) {
        fill_sent_k<double><<<g, 256, 0, c->stream>>>(p.out64.as<double>(), nslots);
        HALA_CHECK_LAUNCH();
    }
    if (p.out32.p) {
        fill_sent_k<float><<<g, 256, 0, c->stream>>>(p.out32.as<float>(), nslots);
        HALA_CHECK_LAUNCH();
    }
}

// ---------------------------------------------------------------------------
// values
// ---------------------------------------------------------------------------
template <class T>
__global__ void gather_vals(int64_t nent, const int *__restrict__ src, const double *__restrict__ f,
                            T *__restrict__ out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    int s = src[e];
    out[e] = (s >= 0) ? (T)f[s] : T(0);
}

template <class T>
static void load_values(hala_ctx c, tri_plan &p, const double *factor, dbuf &ent, dbuf &dg) {
    int64_t nslots = (int64_t)p.nchunk * WAVE;
    ent.reserve((size_t)(p.nent > 0 ? p.nent : 1) * sizeof(T));
    dg.reserve((size_t)(nslots + SLOT_PAD) * sizeof(T));
    if (p.nent > 0) {
        gather_vals<T><<<(int)((p.nent + 255) / 256), 256, 0, c->stream>>>(p.nent, p.ent_src.as<int>(), factor,
                                                                          ent.as<T>());
        HALA_CHECK_LAUNCH();
    }
    gather_vals<T><<<(int)((nslots + 255) / 256), 256, 0, c->stream>>>(nslots, p.dg_src.as<int>(), factor,
                                                                      dg.as<T>());
    HALA_CHECK_LAUNCH();
}

void load_values64(hala_ctx c, tri_plan &p, const double *factor) {
    load_values<double>(c, p, factor, p.ent_val64, p.dg_val64);
    col_load_values(c, p, factor, false);
    p.has64 = true;
}
void load_values32(hala_ctx c, tri_plan &p, const double *factor) {
    load_values<float>(c, p, factor, p.ent_val32, p.dg_val32);
    col_load_values(c, p, factor, true);
    if (!p.out32.p) {
        p.out32.reserve((size_t)p.nchunk * WAVE * sizeof(float));
        tri_rearm(c, p);
    }
    p.has32 = true;
}

}  // namespace hala_amd
