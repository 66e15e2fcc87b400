// ilu.hip -- ILU(0) on the device: symbolic analysis, numeric factorisation
// level by level, and the triangular applies (replacing rocsparse_?csrilu0
// and rocsparse_?csrsv of gpu/hala_gpu_ilu.hpp:81-172).
#include <algorithm>
#include <climits>
#include <cmath>
#include <exception>
#include <future>
#include <thread>

#include "ilu.hpp"

namespace hala_amd {

// diag[i] = first position in row i with column >= i (sparse/hala_sparse_core.hpp:45-60)
__global__ void diag_kernel(int n, const int *__restrict__ pntr, const int *__restrict__ indx, int *__restrict__ diag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int d = pntr[i];
    while (d < pntr[i + 1] && indx[d] < i) d++;
    diag[i] = d;
}

// IKJ ILU(0), level by level: for each strictly lower entry (j,i) of row j in
// column order, a_ji /= u_ii, then a_jk -= a_ji * u_ik over the columns k > i
// both rows hold (merge of the sorted rows).  Rows i of earlier levels are
// final, so each entry sees exactly the operation sequence of the reference's
// KIJ loop (sparse/hala_sparse_utils.hpp:229-254).
//
// Symbolic part of the elimination, once per pattern: for every strictly lower
// entry jc = (j,i) of row j, the position of u_ii and the (ik, jk) position
// pairs its update touches (columns k > i both rows hold, in merge order).
__global__ void ilu_pair_count(int n, const int *__restrict__ pntr, const int *__restrict__ indx,
                               const int *__restrict__ diag, int *__restrict__ cnt) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int je = pntr[j + 1], dj = diag[j];
    for (int jc = pntr[j]; jc < je; jc++) {
        int m = 0;
        if (jc < dj) {
            const int i = indx[jc], ie = pntr[i + 1];
            int ik = diag[i] + 1, jk = jc + 1;
            while (ik < ie && jk < je) {
                const int ci = indx[ik], cj = indx[jk];
                m += ci == cj;
                ik += ci <= cj;
                jk += cj <= ci;
            }
        }
        cnt[jc] = m;
    }
}

__global__ void ilu_pair_fill(int n, const int *__restrict__ pntr, const int *__restrict__ indx,
                              const int *__restrict__ diag, const int *__restrict__ pptr, int *__restrict__ dpos,
                              int2 *__restrict__ pairs) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int je = pntr[j + 1], dj = diag[j];
    for (int jc = pntr[j]; jc < dj; jc++) {
        const int i = indx[jc], ie = pntr[i + 1];
        dpos[jc] = diag[i];
        int ik = diag[i] + 1, jk = jc + 1, o = pptr[jc];
        while (ik < ie && jk < je) {
            const int ci = indx[ik], cj = indx[jk];
            if (ci == cj) pairs[o++] = make_int2(ik, jk);
            ik += ci <= cj;
            jk += cj <= ci;
        }
    }
}

// Numeric ILU(0) in one launch, sync-free.  Wavefront t factors the t-th row
// of the L schedule's level order: for each strictly lower entry jc = (j,i)
// of row j in column order, l = a_ji / u_ii, a_ji = l, then a_jk -= l * u_ik
// over the (ik, jk) pairs of the symbolic pass -- every entry sees the
// operation sequence of the reference's KIJ loop
// (sparse/hala_sparse_utils.hpp:229-254).
//
// Operands from earlier rows (the pivots u_ii and the entries u_ik) are read
// in f, which holds sentinels until the row owning an entry stores its final
// value with an agent-scope store: each value is its own ready flag, so no
// fence or per-row flag is needed.  A row depends only on rows of earlier
// levels, i.e. on lower wavefront indices, and workgroups are dispatched in
// index order on each XCD (as the reductions' collector also assumes): the
// lowest unfinished row has every dependency stored and a running wavefront,
// so every wait ends.  Waits are bounded and raise the wait-error flag (bit
// 1) if they give up.  (Tickets handed out by an atomic would not need the
// dispatch order, but one per row cost ~13 ns each, more than the whole
// factorisation of a 128^3 operator.)
//
// A row is staged in LDS and all its operands from earlier rows are polled up
// front, so a row costs a few memory round trips; the elimination stays in
// order, and the updates of one entry touch distinct positions, so its lanes
// apply them at once.  Rows over the staged limits run the same loop on lane
// 0, in `work` (same indexing as f).
constexpr int ILU_RB = 128;  // row entries staged
constexpr int ILU_PS = 4;    // pairs per lane (256 per row)

__device__ __forceinline__ double poll_factor(const double *p, bool &bad) {
    double v = ld_agent(p);
    for (unsigned spins = 0; is_sent(v); spins++) {
        if (spins > SPIN_LIMIT) {
            bad = true;
            return 0.0;
        }
        __builtin_amdgcn_s_sleep(1);
        v = ld_agent(p);
    }
    return v;
}

__global__ __launch_bounds__(256) void ilu_rows_syncfree(const int4 *__restrict__ frow, const int *__restrict__ fnp,
                                                         int n, const int *__restrict__ dpos,
                                                         const int *__restrict__ pptr, const int2 *__restrict__ pairs,
                                                         const double *__restrict__ vals, double *f, double *work,
                                                         int *err) {
    __shared__ double rowbuf[4][ILU_RB];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & (WAVE - 1);
    const int t = blockIdx.x * 4 + w;
    if (t >= n) return;  // a whole wavefront
    double *rb = rowbuf[w];
    bool bad = false;
    const int4 ri = frow[t];
    const int rs = ri.x, re = ri.y, dj = ri.z, p0 = ri.w, np = fnp[t];
    const int len = re - rs, nl = dj - rs;
    if (nl == 0) {  // no strictly lower entry: the row is final as given
        for (int e = lane; e < len; e += WAVE) st_agent(f + rs + e, vals[rs + e]);
        return;
    }
    if (len > ILU_RB || nl >= WAVE || np > ILU_PS * WAVE) {
        if (lane == 0) {
            for (int e = rs; e < re; e++) work[e] = vals[e];
            for (int jc = rs; jc < dj; jc++) {
                const double l = work[jc] / poll_factor(f + dpos[jc], bad);
                work[jc] = l;
                for (int p = pptr[jc]; p < pptr[jc + 1]; p++) {
                    const int2 q = pairs[p];
                    work[q.y] = work[q.y] - l * poll_factor(f + q.x, bad);
                }
            }
            for (int e = rs; e < re; e++) st_agent(f + e, work[e]);
            if (bad) atomicOr(err, 1);
        }
        return;
    }
    for (int e = lane; e < len; e += WAVE) rb[e] = vals[rs + e];
    const int pp = lane <= nl ? pptr[rs + lane] : 0;  // pair range of lower entry `lane`
    // operands from earlier rows: bit 0 the pivot of lower entry `lane`, bit k+1 pair k*64+lane
    const int pi = lane < nl ? dpos[rs + lane] : 0;
    unsigned pend = lane < nl ? 1u : 0u;
    int y[ILU_PS], xi[ILU_PS];
    double fx[ILU_PS];
#pragma unroll
    for (int k = 0; k < ILU_PS; k++) {
        const int idx = k * WAVE + lane;
        y[k] = 0;
        xi[k] = 0;
        fx[k] = 0.0;
        if (idx < np) {
            const int2 q = pairs[p0 + idx];
            y[k] = q.y - rs;
            xi[k] = q.x;
            pend |= 2u << k;
        }
    }
    double piv = 1.0;
    for (unsigned spins = 0;; spins++) {
        const double v0 = (pend & 1u) ? ld_agent(f + pi) : 1.0;
        double v[ILU_PS];
#pragma unroll
        for (int k = 0; k < ILU_PS; k++) v[k] = (pend >> (k + 1) & 1u) ? ld_agent(f + xi[k]) : 0.0;
        if ((pend & 1u) && !is_sent(v0)) {
            piv = v0;
            pend &= ~1u;
        }
#pragma unroll
        for (int k = 0; k < ILU_PS; k++)
            if ((pend >> (k + 1) & 1u) && !is_sent(v[k])) {
                fx[k] = v[k];
                pend &= ~(2u << k);
            }
        if (!__any(pend != 0u)) break;
        if (spins > SPIN_LIMIT) {
            bad = bad || pend != 0u;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    for (int jc = 0; jc < nl; jc++) {
        const int pb = __shfl(pp, jc, WAVE) - p0, pe = __shfl(pp, jc + 1, WAVE) - p0;
        const double l = rb[jc] / __shfl(piv, jc, WAVE);
        if (lane == 0) rb[jc] = l;
#pragma unroll
        for (int k = 0; k < ILU_PS; k++) {
            const int idx = k * WAVE + lane;
            if (idx >= pb && idx < pe) rb[y[k]] = rb[y[k]] - l * fx[k];
        }
    }
    for (int e = lane; e < len; e += WAVE) st_agent(f + rs + e, rb[e]);
    if (__any(bad) && lane == 0) atomicOr(err, 1);
}

// every factor entry to the sentinel (not yet stored by its row)
__global__ void factor_arm(double *__restrict__ f, int64_t nnz) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
        f[e] = sentinel64();
}

// the factorisation's row records, in the level order of the L schedule
__global__ void ilu_row_info(int n, const int *__restrict__ order, const int *__restrict__ pntr,
                             const int *__restrict__ diag, const int *__restrict__ pptr, int4 *__restrict__ frow,
                             int *__restrict__ fnp) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int j = order[t], rs = pntr[j], dj = diag[j];
    frow[t] = make_int4(rs, pntr[j + 1], dj, pptr[rs]);
    fnp[t] = pptr[dj] - pptr[rs];
}

__global__ void pivot_check(int n, const int *__restrict__ diag, const double *__restrict__ f, int *bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u = f[diag[i]];
    if (u == 0.0 || !isfinite(u)) atomicMin(bad, i);
}

__global__ void set_int(int *p, int v) { *p = v; }

// the pencil schedule applies (a stencil pattern, and the option allows it)
static bool use_stencil(hala_ilu M) { return M->sp && M->ctx->opt.tri_stencil; }

bool ilu_fused64(hala_ilu M) {
    return M->factored && !use_stencil(M) && col_usable(M->ctx, M->L, true) && col_usable(M->ctx, M->U, true) &&
           M->L.col.hand64;
}

void ilu_apply64_mid(hala_ilu M, const int *gate, const col_axpy *ax) {
    hala_ctx c = M->ctx;
    // L's helpers re-arm U's polled output, U's helpers L's for the next apply
    col_sweep64(c, M->L, TRI_SUB, 1.0, gate, &M->U, &M->U, ax);
    col_sweep64(c, M->U, TRI_SUB, 1.0, gate, nullptr, &M->L);
}

void ilu_apply64(hala_ilu M, const double *x, double *r, const int *gate) {
    hala_ctx c = M->ctx;
    HALA_REQUIRE(M->factored, HALA_ERR_ARG, "ILU apply before factorisation");
    if (use_stencil(M)) {  // natural order in and out, no hand-off
        stencil_apply<double, double, double>(M, x, r, gate);
        return;
    }
    if (ilu_fused64(M)) {  // L's output handed to U in slot order
        tri_col_exec<double, double>(c, M->L, TRI_SUB, 1.0, x, nullptr, gate, &M->U);
        col_sweep64(c, M->U, TRI_SUB, 1.0, gate, nullptr, &M->L);
        col_scatter64(c, M->U, r, gate);
        return;
    }
    tri_exec<double, double, double>(c, M->L, TRI_SUB, 1.0, x, M->tmp64.as<double>(), gate);
    tri_exec<double, double, double>(c, M->U, TRI_SUB, 1.0, M->tmp64.as<double>(), r, gate);
}

void ilu_apply32x64(hala_ilu M, const double *x, double *r, const int *gate) {
    hala_ctx c = M->ctx;
    HALA_REQUIRE(M->factored && M->has32, HALA_ERR_ARG, "FP32 ILU apply needs hala_ilu_factor_f64(..., want_f32=1)");
    if (c->opt.ilu32_jacobi > 0) {
        ilu_apply32x64_jacobi(M, x, r, gate, c->opt.ilu32_jacobi);
        return;
    }
    if (use_stencil(M)) {
        stencil_apply<float, double, double>(M, x, r, gate);
        return;
    }
    tri_exec<float, double, double>(c, M->L, TRI_SUB, 1.0f, x, M->tmp64.as<double>(), gate);
    tri_exec<float, double, double>(c, M->U, TRI_SUB, 1.0f, M->tmp64.as<double>(), r, gate);
}

void ilu_prepare_adjoint(hala_ilu M) {
    hala_ctx c = M->ctx;
    if (!M->adjoint) {
        if (!M->tp.built) build_transpose(c, M->n, M->n, M->nnz, M->pntr, M->indx, M->tp);
        // U^T w = x: row k of U^T lists U_ik, i < k, ascending i; diagonal U_kk
        build_deps(c, M->Ut, M->n, DEP_T_LOWER, M->pntr, M->indx, M->diag.as<int>(), &M->tp, false, false);
        build_schedule(c, M->Ut, false);
        // L^T y = w: row k of L^T lists L_ik, i > k, descending i; unit diagonal
        build_deps(c, M->Lt, M->n, DEP_T_UPPER, M->pntr, M->indx, M->diag.as<int>(), &M->tp, true, true);
        build_schedule(c, M->Lt, false);
        M->adjoint = true;
    }
    if (!M->adjoint_loaded) {
        load_values64(c, M->Ut, M->factor.as<double>());
        load_values64(c, M->Lt, M->factor.as<double>());
        M->adjoint_loaded = true;
    }
}

void ilu_apply_t64(hala_ilu M, const double *x, double *r, const int *gate) {
    hala_ctx c = M->ctx;
    HALA_REQUIRE(M->factored, HALA_ERR_ARG, "ILU apply before factorisation");
    ilu_prepare_adjoint(M);
    tri_exec<double, double, double>(c, M->Ut, TRI_SUB, 1.0, x, M->tmp64.as<double>(), gate);
    tri_exec<double, double, double>(c, M->Lt, TRI_SUB, 1.0, M->tmp64.as<double>(), r, gate);
}

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

static void ilu_pairs(hala_ctx c, hala_ilu M) {
    const int n = M->n, g = (n + 255) / 256;
    const int64_t nnz = M->nnz;
    tmp_scope ts(c);
    dbuf &cnt = ts.get();
    cnt.reserve((size_t)nnz * sizeof(int));
    M->pptr.reserve((size_t)(nnz + 1) * sizeof(int));
    M->dpos.reserve((size_t)nnz * sizeof(int));
    ilu_pair_count<<<g, 256, 0, c->stream>>>(n, M->pntr, M->indx, M->diag.as<int>(), cnt.as<int>());
    HALA_CHECK_LAUNCH();
    exclusive_scan_i32(c, cnt.as<int>(), M->pptr.as<int>(), (int)nnz);
    const int np = fetch_i32(c, M->pptr.as<int>() + nnz);
    M->pairs.reserve((size_t)std::max(np, 1) * sizeof(int2));
    ilu_pair_fill<<<g, 256, 0, c->stream>>>(n, M->pntr, M->indx, M->diag.as<int>(), M->pptr.as<int>(),
                                            M->dpos.as<int>(), M->pairs.as<int2>());
    HALA_CHECK_LAUNCH();
}

// rows in L's level order with their ranges (after ilu_pairs and L's schedule)
static void ilu_rows(hala_ctx c, hala_ilu M) {
    const int n = M->n, g = (n + 255) / 256;
    // a copy of the level order: the plan's own may be re-sorted later
    M->frow.reserve((size_t)n * sizeof(int4));
    M->fnp.reserve((size_t)n * sizeof(int));
    ilu_row_info<<<g, 256, 0, c->stream>>>(n, M->L.perm.as<int>(), M->pntr, M->diag.as<int>(), M->pptr.as<int>(),
                                           M->frow.as<int4>(), M->fnp.as<int>());
    HALA_CHECK_LAUNCH();
}

int hala_ilu_create(hala_ctx c, int n, int64_t nnz, const int *pntr, const int *indx, hala_ilu *out) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(c && out, HALA_ERR_ARG, "hala_ilu_create: null argument");
        *out = nullptr;
        HALA_REQUIRE(n > 0 && nnz > 0 && nnz < INT_MAX && pntr && indx, HALA_ERR_ARG, "hala_ilu_create: bad sizes");
        set_device(c);
        int bad = validate_pattern(c, n, n, nnz, pntr, indx, 1);
        HALA_REQUIRE(bad < 0, HALA_ERR_PATTERN,
                     "hala_ilu_create: row " + std::to_string(bad) +
                         " has unsorted/out-of-range columns or no diagonal entry");
        auto *M = new hala_ilu_s();
        // the factor's buffers come from the context's cache; hala_ilu_destroy
        // synchronises the context's stream before returning them
        alloc_scope as(c->cache.get());
        try {
            M->ctx = c;
            M->n = n;
            M->nnz = nnz;
            M->pntr = pntr;
            M->indx = indx;
            M->diag.reserve((size_t)n * sizeof(int));
            diag_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(n, pntr, indx, M->diag.as<int>());
            HALA_CHECK_LAUNCH();
            // The two analyses are independent and mostly latency-bound: U's runs
            // on the auxiliary context's stream, from a second host thread.
            // U's levels are L's mirrored where that passes U's own level
            // recurrence (one pass instead of a second level-set kernel, which
            // would have to wait for L's: each takes the whole chip).
            hala_ctx a = aux_ctx(c);
            hipEvent_t ready = nullptr, lev_ready = nullptr;
            HALA_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
            if (hipEventCreateWithFlags(&lev_ready, hipEventDisableTiming) != hipSuccess) {
                (void)hipEventDestroy(ready);
                throw status_error{HALA_ERR_HIP, "hala_ilu_create: event creation failed"};
            }
            std::promise<int> lnlev;  // L's level count once its levels are on the stream (0: none)
            std::future<int> lnlev_f = lnlev.get_future();
            bool lnlev_set = false;
            std::exception_ptr uerr;
            try {
                HALA_HIP(hipEventRecord(ready, c->stream));
                HALA_HIP(hipStreamWaitEvent(a->stream, ready, 0));
                std::thread ut([&] {
                    try {
                        set_device(a);
                        alloc_scope aas(a->cache.get());  // blocks reused on this stream only
                        // U's dependencies before L's level-set kernel takes the chip, then the
                        // elimination pairs beside it: this thread waits for L's levels anyway
                        build_deps(a, M->U, n, DEP_UPPER, pntr, indx, M->diag.as<int>(), nullptr, false, true);
                        ilu_pairs(a, M);
                        const int nl = lnlev_f.get();
                        bool mirrored = false;
                        if (nl > 0) {
                            HALA_HIP(hipStreamWaitEvent(a->stream, lev_ready, 0));
                            mirrored = mirror_levels(a, M->U, M->L.lev.as<int>(), nl);
                        }
                        if (!mirrored) compute_levels(a, M->U);
                        build_schedule(a, M->U, false, true);
                        HALA_HIP(hipStreamSynchronize(a->stream));
                    } catch (...) {
                        uerr = std::current_exception();
                    }
                });
                try {
                    build_deps(c, M->L, n, DEP_LOWER, pntr, indx, M->diag.as<int>(), nullptr, true, false);
                    compute_levels(c, M->L);
                    HALA_HIP(hipEventRecord(lev_ready, c->stream));
                    lnlev.set_value(M->L.nlev);
                    lnlev_set = true;
                    build_schedule(c, M->L, false, true);
                } catch (...) {
                    if (!lnlev_set) lnlev.set_value(0);
                    ut.join();
                    throw;
                }
                ut.join();
                if (!uerr) ilu_rows(c, M);  // needs the pair pointers (the other stream is idle after the join)
            } catch (...) {
                (void)hipEventDestroy(ready);
                (void)hipEventDestroy(lev_ready);
                throw;
            }
            (void)hipEventDestroy(ready);
            (void)hipEventDestroy(lev_ready);
            if (uerr) std::rethrow_exception(uerr);
            M->factor.reserve((size_t)nnz * sizeof(double));
            M->tmp64.reserve((size_t)n * sizeof(double));
            // with option tri_stencil, a stencil pattern also gets the pencil schedule of its sweeps
            if (c->opt.tri_stencil && !stencil_build(M)) M->sp.reset();
        } catch (...) {
            delete M;
            throw;
        }
        *out = M;
    });
}

int hala_ilu_destroy(hala_ilu M) {
    return guard([&] {
        if (!M) return;
        (void)hipSetDevice(M->ctx->device);
        (void)hipStreamSynchronize(M->ctx->stream);
        delete M;
    });
}

int hala_ilu_info(hala_ilu M, hala_ilu_info_t *info) {
    return guard([&] {
        HALA_REQUIRE(M && info, HALA_ERR_ARG, "hala_ilu_info: null argument");
        info->n = M->n;
        info->nnz = M->nnz;
        info->nlev_lower = M->L.nlev;
        info->nlev_upper = M->U.nlev;
        info->nchunk_lower = M->L.nchunk;
        info->nchunk_upper = M->U.nchunk;
        info->slots_lower = M->L.nent;
        info->slots_upper = M->U.nent;
    });
}

int hala_ilu_stencil_info(hala_ilu M, int *out) {
    return guard([&] {
        HALA_REQUIRE(M && out, HALA_ERR_ARG, "hala_ilu_stencil_info: null argument");
        const stencil_plan *sp = M->sp.get();
        out[0] = sp ? sp->kind : 0;
        out[1] = sp ? sp->nx : 0;
        out[2] = sp ? sp->ny : 0;
        out[3] = sp ? sp->nz : 0;
        out[4] = use_stencil(M) ? 1 : 0;
    });
}

int hala_ilu_sweep_info(hala_ilu M, int upper, int *out) {
    return guard([&] {
        HALA_REQUIRE(M && out, HALA_ERR_ARG, "hala_ilu_sweep_info: null argument");
        const tri_colplan &q = upper ? M->U.col : M->L.col;
        out[0] = q.ok ? 1 : 0;
        out[1] = q.ncol;
        out[2] = q.CP;
        out[3] = q.K;
        out[4] = q.maxext;
        out[5] = q.why;
        out[6] = q.split;
        out[7] = q.plane;
    });
}

int hala_ilu_factor_f64(hala_ilu M, const double *vals, int want_f32) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(M && vals, HALA_ERR_ARG, "hala_ilu_factor_f64: null argument");
        hala_ctx c = M->ctx;
        set_device(c);
        alloc_scope as(c->cache.get());
        double *f = M->factor.as<double>();
        M->fwork.reserve((size_t)M->nnz * sizeof(double));
        factor_arm<<<grid_for(M->nnz, 256 * 4, 8192), 256, 0, c->stream>>>(f, M->nnz);
        ilu_rows_syncfree<<<(M->n + 3) / 4, 256, 0, c->stream>>>(M->frow.as<int4>(), M->fnp.as<int>(), M->n,
                                                                M->dpos.as<int>(), M->pptr.as<int>(),
                                                                M->pairs.as<int2>(), vals, f, M->fwork.as<double>(),
                                                                wait_error_flag(c));
        HALA_CHECK_LAUNCH();
        int *bad = reinterpret_cast<int *>(c->scal.as<double>() + S_DOT + 6);
        set_int<<<1, 1, 0, c->stream>>>(bad, INT_MAX);
        pivot_check<<<(M->n + 255) / 256, 256, 0, c->stream>>>(M->n, M->diag.as<int>(), f, bad);
        HALA_CHECK_LAUNCH();
        M->factored = false;
        M->adjoint_loaded = false;
        M->jac_values = false;
        M->jac_values64 = false;
        // The values are packed without waiting for the host; the
        // factorisation's dependency waits and its pivots are checked once
        // everything has run (packing a factor that fails is harmless: it
        // stays unfactored).
        // the two triangles' values are packed side by side: U's on the
        // auxiliary context's stream, joined back before anything reads them
        // (scattered gathers from the factor, mostly latency)
        hala_ctx a = aux_ctx(c);
        hipEvent_t ev[2] = {nullptr, nullptr};
        HALA_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        if (hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
            (void)hipEventDestroy(ev[0]);
            throw status_error{HALA_ERR_HIP, "hala_ilu_factor_f64: event creation failed"};
        }
        try {
            HALA_HIP(hipEventRecord(ev[0], c->stream));
            HALA_HIP(hipStreamWaitEvent(a->stream, ev[0], 0));
            load_values64(a, M->U, f);
            if (want_f32) load_values32(a, M->U, f);
            HALA_HIP(hipEventRecord(ev[1], a->stream));
            load_values64(c, M->L, f);
            if (want_f32) load_values32(c, M->L, f);
            HALA_HIP(hipStreamWaitEvent(c->stream, ev[1], 0));
        } catch (...) {
            (void)hipStreamSynchronize(a->stream);
            (void)hipEventDestroy(ev[0]);
            (void)hipEventDestroy(ev[1]);
            throw;
        }
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        col_set_handoff(c, M->L, M->U);
        if (M->sp) stencil_load(M, want_f32);
        check_wait_error(c);  // synchronises; a dependency wait that gave up throws here
        const int b = fetch_i32(c, bad);
        HALA_REQUIRE(b == INT_MAX, HALA_ERR_PIVOT,
                     "hala_ilu_factor_f64: zero or non-finite pivot in row " + std::to_string(b));
        if (want_f32) M->has32 = true;
        M->factored = true;
    });
}

int hala_ilu_apply_f64(hala_ilu M, const double *x, double *r) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(M && x && r, HALA_ERR_ARG, "hala_ilu_apply_f64: null argument");
        set_device(M->ctx);
        HALA_REQUIRE(M->factored, HALA_ERR_ARG, "ILU apply before factorisation");
        if (M->ctx->opt.ilu64_jacobi > 0) ilu_apply64_jacobi(M, x, r, nullptr, M->ctx->opt.ilu64_jacobi);
        else ilu_apply64(M, x, r);
        if (M->ctx->opt.sync_checks) check_wait_error(M->ctx);
    });
}

int hala_ilu_apply_f32(hala_ilu M, const float *x, float *r) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(M && x && r, HALA_ERR_ARG, "hala_ilu_apply_f32: null argument");
        HALA_REQUIRE(M->factored && M->has32, HALA_ERR_ARG, "FP32 ILU apply needs want_f32 at factorisation");
        hala_ctx c = M->ctx;
        set_device(c);
        if (use_stencil(M)) {
            stencil_apply<float, float, float>(M, x, r, nullptr);
        } else {
            M->tmp32.reserve((size_t)M->n * sizeof(float));
            tri_exec<float, float, float>(c, M->L, TRI_SUB, 1.0f, x, M->tmp32.as<float>());
            tri_exec<float, float, float>(c, M->U, TRI_SUB, 1.0f, M->tmp32.as<float>(), r);
        }
        if (c->opt.sync_checks) check_wait_error(c);
    });
}

int hala_ilu_apply_f32x64(hala_ilu M, const double *x, double *r) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(M && x && r, HALA_ERR_ARG, "hala_ilu_apply_f32x64: null argument");
        set_device(M->ctx);
        ilu_apply32x64(M, x, r);
        if (M->ctx->opt.sync_checks) check_wait_error(M->ctx);
    });
}

int hala_ilu_apply_t_f64(hala_ilu M, const double *x, double *r) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(M && x && r, HALA_ERR_ARG, "hala_ilu_apply_t_f64: null argument");
        set_device(M->ctx);
        ilu_apply_t64(M, x, r);
        if (M->ctx->opt.sync_checks) check_wait_error(M->ctx);
    });
}

int hala_ilu_get_symbolic(hala_ilu M, int upper, int *diag, int *lev, int *level_ptr, int *perm) {
    return guard([&] {
        HALA_REQUIRE(M, HALA_ERR_ARG, "hala_ilu_get_symbolic: null handle");
        hala_ctx c = M->ctx;
        set_device(c);
        tri_plan &p = upper ? M->U : M->L;
        if (perm) sort_perm(c, p);
        auto cp = [&](int *dst, const dbuf &src, size_t count) {
            if (dst) HALA_HIP(hipMemcpyAsync(dst, src.p, count * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
        };
        cp(diag, M->diag, M->n);
        cp(lev, p.lev, M->n);
        cp(level_ptr, p.level_ptr, (size_t)p.nlev + 1);
        cp(perm, p.perm, M->n);
        HALA_HIP(hipStreamSynchronize(c->stream));
    });
}

int hala_ilu_get_factor_f64(hala_ilu M, double *factor) {
    return guard([&] {
        HALA_REQUIRE(M && factor, HALA_ERR_ARG, "hala_ilu_get_factor_f64: null argument");
        hala_ctx c = M->ctx;
        set_device(c);
        HALA_HIP(hipMemcpyAsync(factor, M->factor.p, (size_t)M->nnz * sizeof(double), hipMemcpyDeviceToDevice,
                                c->stream));
        HALA_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
