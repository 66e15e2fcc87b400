// cg.hip -- preconditioned conjugate gradient, resident on the device.
//
// Step order and scalars follow HALA's solve_cg_core
// (hex/solvers/hala_solvers_cg.hpp:92-156 with the lambdas of :203-226):
//   r = b - A x; z = M r; p = z; zr = <r,z>; it = 1
//   loop: Ap = A p; it++; a = zr / <p,Ap>; x += a p; r += (-a) Ap;
//         stop if it == max_iter or norm2(r) < tol;
//         z = M r; nzr = <r,z>; b = nzr / zr; p = b p; p += 1*z; zr = nzr
// Every scalar lives on the device; each kernel returns at once when the
// stop flag is set, so the host launches passes in batches and reads the
// flag once per batch instead of blocking on every dot product.
#include <algorithm>
#include <cmath>

#include "blas1.hpp"
#include "solvers.hpp"

namespace hala_amd {

constexpr int EK = 2;  // canonical chunks per block in the elementwise kernels (128^3 PCG: 1, 2, 4 -> 79.06, 78.85, 79.27 ms)

enum { PREC_ID = 0, PREC_ILU = 1, PREC_JAC = 2, PREC_ILU_T = 3 };

__global__ void cg_init_kernel(cg_state *st, double tol, int max_iter, double *hist, int max_hist) {
    st->zr = st->alpha = st->beta = st->rn = 0.0;
    st->tol = tol;
    st->hist = hist;
    st->it = 1;
    st->done = 0;
    st->max_iter = max_iter;
    st->max_hist = max_hist;
    st->status = 0;
}

// r = b + (-1)*(1*(A x) + 0): HALA's p = A x (beta 0) then axpy(-1, p, r) with r = b.
template <bool TRANS>
__global__ __launch_bounds__(256) void cg_residual_kernel(int n, const int *__restrict__ pntr,
                                                          const int *__restrict__ indx, const double *__restrict__ vals,
                                                          const double *__restrict__ x, const double *__restrict__ b,
                                                          double *__restrict__ r) {
    __shared__ double prod[spmv_tile<double>::value];
    double acc = csr_block_rowsum<double, double, double>(n, pntr, indx, vals, x, prod);
    int i = blockIdx.x * SPMV_ROWS + threadIdx.x;
    if (i < n) r[i] = b[i] + (-1.0) * (1.0 * acc + 0.0);
}

// p = z, zr = <r, z>; z = r (identity) or r ./ d (Jacobi) formed on the fly.
template <int PREC>
__global__ __launch_bounds__(256) void cg_start_kernel(int n, const double *__restrict__ r,
                                                       const double *__restrict__ z, const double *__restrict__ d,
                                                       double *__restrict__ p, double *part, int *err,
                                                       cg_state *st) {
    __shared__ double sh[CB * EK];
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    double v[EK];
#pragma unroll
    for (int k = 0; k < EK; k++) {
        int64_t i = ((int64_t)blockIdx.x * EK + k) * CB + threadIdx.x;
        double a = 0.0;
        if (i < n) {
            double ri = r[i];
            double zi = (PREC == PREC_ID) ? ri : (PREC == PREC_JAC ? ri / d[i] : z[i]);
            p[i] = zi;
            a = ri * zi;
        }
        v[k] = a;
    }
    tree256<double, EK>(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < EK; k++) {
            int64_t b = (int64_t)blockIdx.x * EK + k;
            if (b < nb) st_agent(part + b, v[k]);
        }
    if (is_collector()) {
        double t = final_pass<double>(part, nb, sh, err);
        if (threadIdx.x == 0) st->zr = t;
    }
}

// The 256-row block group a workgroup handles: with xcd, the XCD-aware
// virtual block (every XCD takes a contiguous share of the rows, so that the
// scattered stores of neighbouring rows into one line of a slot order merge
// in one L2 instead of leaving the line partially written by several; the
// grid is then xcd_grid(nvirt) and workgroups past nvirt have nothing to do),
// else its own index.  Workgroup 0 handles block 0 either way.  (128^3
// ILU-PCG: the update with its scatter into L's slot order 78.6 -> 77.6 ms
// per solve; the z gather from U's slot order and the fused product were
// slower in this order.)
__device__ __forceinline__ int cg_block(int xcd, int nvirt) { return xcd ? xcd_block(nvirt) : (int)blockIdx.x; }

// Ap = A p (beta 0) fused with <p, Ap>; the last block forms alpha = zr / <p,Ap>.
// One block = 256 rows = one canonical partial.
template <bool NT, class TI = int>
__global__ __launch_bounds__(256) void cg_spmv_kernel(int n, const int *__restrict__ pntr,
                                                      const TI *__restrict__ indx, const double *__restrict__ vals,
                                                      const double *__restrict__ p, double *__restrict__ ap,
                                                      double *part, int *err, cg_state *st) {
    if (st->done) return;
    __shared__ double prod[spmv_tile<double>::value];
    double *sh = prod;  // LDS reused for the tree
    const int nb = (n + SPMV_ROWS - 1) / SPMV_ROWS;
    // grid-stride over 256-row blocks: one arrival (atomic) per workgroup, not per row block
    for (int blk = blockIdx.x; blk < nb; blk += gridDim.x) {
        double acc = csr_block_rowsum<double, double, double, false, NT, TI>(n, pntr, indx, vals, p, prod, blk);
        int i = blk * SPMV_ROWS + threadIdx.x;
        double v = 0.0;
        if (i < n) {
            double y = 1.0 * acc + 0.0;
            ap[i] = y;
            v = p[i] * y;
        }
        v = tree256_1<double>(v, sh);
        if (threadIdx.x == 0) st_agent(part + blk, v);
    }
    if (is_collector()) {
        double t = final_pass<double>(part, nb, sh, err);
        if (threadIdx.x == 0) {
            st->alpha = st->zr / t;
            st->it = st->it + 1;
        }
    }
}

// Identity / Jacobi CG on small systems, where every launch costs more than
// its work: the direction update p = beta p + z of the previous pass is formed
// inside the product, for each gathered entry and for the block's own rows
// (z = r or r ./ d on the fly; the same two roundings as cg_pupdate_kernel),
// from pold into pnew (the next pass swaps them).  Then Ap = A pnew and
// <pnew, Ap> as in cg_spmv_kernel: each row summed in CSR order, a canonical
// partial per 256-row block, alpha from the collector.
constexpr int PZ_MAX_ROWS = 1 << 18;  // above this the separate update's streaming access pays
template <int PREC>
__global__ __launch_bounds__(256) void cg_spmv_pz_kernel(int n, const int *__restrict__ pntr,
                                                         const int *__restrict__ indx, const double *__restrict__ vals,
                                                         const double *__restrict__ pold, const double *__restrict__ r,
                                                         const double *__restrict__ d, double *__restrict__ pnew,
                                                         double *__restrict__ ap, double *part, int *err,
                                                         cg_state *st) {
    if (st->done) return;
    __shared__ double sh[CB];
    const double b = st->beta;
    auto pz = [&](int j) {
        const double zj = (PREC == PREC_ID) ? r[j] : r[j] / d[j];
        return b * pold[j] + 1.0 * zj;
    };
    const int nb = (n + CB - 1) / CB;
    const int i = blockIdx.x * CB + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        double acc = 0.0;
        const int e1 = pntr[i + 1];
        for (int e = pntr[i]; e < e1; e++) acc = acc + vals[e] * pz(indx[e]);
        const double y = 1.0 * acc + 0.0;
        const double pi = pz(i);
        pnew[i] = pi;
        ap[i] = y;
        v = pi * y;
    }
    v = tree256_1<double>(v, sh);
    if (threadIdx.x == 0) st_agent(part + blockIdx.x, v);
    if (is_collector()) {
        double t = final_pass<double>(part, nb, sh, err);
        if (threadIdx.x == 0) {
            st->alpha = st->zr / t;
            st->it = st->it + 1;
        }
    }
}

// x += a p; r += (-a) Ap; <r,r> (and <r, r./d> for Jacobi).  The last block
// takes the stop decision and, for identity/Jacobi, the new beta.
// lslot (ILU on the column layout, may be null): r also goes straight into the
// L sweep's slot-order input (the sweep's gather pass, fused here), and x is
// left alone (null): the L sweep's helper workgroups update it meanwhile, and
// cg_xfinal does the pass whose sweep the stop gated off.
template <int PREC>
__global__ __launch_bounds__(256) void cg_update_kernel(int n, const double *__restrict__ p,
                                                        const double *__restrict__ ap, const double *__restrict__ d,
                                                        double *__restrict__ x, double *__restrict__ r, double *part,
                                                        double *part2, int *err, cg_state *st,
                                                        const int *__restrict__ lslot = nullptr,
                                                        double *__restrict__ lxs = nullptr, int xcd = 0) {
    if (st->done) return;
    __shared__ double sh[CB * EK * 2];
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    const int vb = cg_block(xcd, (int)((nb + EK - 1) / EK));
    if ((int64_t)vb * EK >= nb) return;
    const double a = st->alpha;
    double v[2 * EK];
#pragma unroll
    for (int k = 0; k < EK; k++) {
        int64_t i = ((int64_t)vb * EK + k) * CB + threadIdx.x;
        double rr = 0.0, rz = 0.0;
        if (i < n) {
            if (x) x[i] = x[i] + a * p[i];
            double ri = r[i] + (-a) * ap[i];
            r[i] = ri;
            if (lslot) lxs[lslot[i]] = ri;
            rr = ri * ri;
            if (PREC == PREC_JAC) rz = ri * (ri / d[i]);
        }
        v[k] = rr;
        v[EK + k] = rz;
    }
    if (PREC == PREC_JAC) {
        tree256<double, 2 * EK>(v, sh);
    } else {
        double w[EK];
#pragma unroll
        for (int k = 0; k < EK; k++) w[k] = v[k];
        tree256<double, EK>(w, sh);
#pragma unroll
        for (int k = 0; k < EK; k++) v[k] = w[k];
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < EK; k++) {
            int64_t b = (int64_t)vb * EK + k;
            if (b < nb) {
                st_agent(part + b, v[k]);
                if (PREC == PREC_JAC) st_agent(part2 + b, v[EK + k]);
            }
        }
    if (is_collector()) {
        double rr = final_pass<double>(part, nb, sh, err);
        double rz = 0.0;
        if (PREC == PREC_JAC) rz = final_pass<double>(part2, nb, sh, err);
        if (threadIdx.x == 0) {
            double rn = sqrt(rr);
            st->rn = rn;
            int it = st->it;
            if (st->hist && it - 2 < st->max_hist) st->hist[it - 2] = rn;
            if (it == st->max_iter || rn < st->tol) {
                st->done = 1;
            } else if (!isfinite(rn)) {
                st->done = 1;
                st->status = 1;
            } else if (PREC == PREC_ID || PREC == PREC_JAC) {
                double nzr = (PREC == PREC_ID) ? rr : rz;
                st->beta = nzr / st->zr;
                st->zr = nzr;
            }
        }
    }
}

// nzr = <r, z>; beta = nzr / zr; zr = nzr   (ILU preconditioner)
__global__ __launch_bounds__(256) void cg_rz_kernel(int n, const double *__restrict__ r,
                                                    const double *__restrict__ z, double *part, int *err,
                                                    cg_state *st) {
    if (st->done) return;
    __shared__ double sh[CB * EK];
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    double v[EK];
#pragma unroll
    for (int k = 0; k < EK; k++) {
        int64_t i = ((int64_t)blockIdx.x * EK + k) * CB + threadIdx.x;
        v[k] = (i < n) ? r[i] * z[i] : 0.0;
    }
    tree256<double, EK>(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < EK; k++) {
            int64_t b = (int64_t)blockIdx.x * EK + k;
            if (b < nb) st_agent(part + b, v[k]);
        }
    if (is_collector()) {
        double t = final_pass<double>(part, nb, sh, err);
        if (threadIdx.x == 0) {
            st->beta = t / st->zr;
            st->zr = t;
        }
    }
}

// z from the U sweep's slot-order output (its scatter pass, fused here) and
// nzr = <r, z> in the canonical order; beta = nzr / zr; zr = nzr
__global__ __launch_bounds__(256) void cg_zscatter_rz_kernel(int n, const int *__restrict__ uslot,
                                                             const double *__restrict__ uout,
                                                             const double *__restrict__ r, double *__restrict__ z,
                                                             double *part, int *err, cg_state *st) {
    if (st->done) return;
    __shared__ double sh[CB * EK];
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    double v[EK];
#pragma unroll
    for (int k = 0; k < EK; k++) {
        int64_t i = ((int64_t)blockIdx.x * EK + k) * CB + threadIdx.x;
        double a = 0.0;
        if (i < n) {
            const double zi = uout[uslot[i]];
            z[i] = zi;
            a = r[i] * zi;
        }
        v[k] = a;
    }
    tree256<double, EK>(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < EK; k++) {
            int64_t b = (int64_t)blockIdx.x * EK + k;
            if (b < nb) st_agent(part + b, v[k]);
        }
    if (is_collector()) {
        double t = final_pass<double>(part, nb, sh, err);
        if (threadIdx.x == 0) {
            st->beta = t / st->zr;
            st->zr = t;
        }
    }
}

// p = beta p; p += 1 z   (z = r for identity, r ./ d for Jacobi)
template <int PREC>
__global__ void cg_pupdate_kernel(int n, const double *__restrict__ z, const double *__restrict__ r,
                                  const double *__restrict__ d, double *__restrict__ p, const cg_state *st) {
    if (st->done) return;
    const double b = st->beta;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        double zi = (PREC == PREC_ID) ? r[i] : (PREC == PREC_JAC ? r[i] / d[i] : z[i]);
        p[i] = b * p[i] + 1.0 * zi;
    }
}

// x += alpha p for the last pass of a fused ILU-PCG solve (whose L sweep, and
// with it the helpers' update of x, the stop flag gated off)
__global__ void cg_xfinal(int n, double *__restrict__ x, const double *__restrict__ p, const cg_state *st) {
    const double a = st->alpha;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = x[i] + a * p[i];
}

__global__ void cg_gate_copy(int n, const int *done, const double *__restrict__ src, double *__restrict__ dst) {
    if (*done) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

void cg_read(hala_ctx c, int *it, int *done, int *status, double *rn) {
    cg_state *st = cg_state_ptr(c);
    HALA_HIP(hipMemcpyAsync(c->host_pinned, st, sizeof(cg_state), hipMemcpyDeviceToHost, c->stream));
    HALA_HIP(hipStreamSynchronize(c->stream));
    cg_state h;
    std::memcpy(&h, c->host_pinned, sizeof(cg_state));
    if (it) *it = h.it;
    if (done) *done = h.done;
    if (status) *status = h.status;
    if (rn) *rn = h.rn;
}

// Operator/preconditioner bundle of one CG run.
struct cg_problem {
    hala_csr A;
    const double *vals;
    int prec;
    hala_ilu M;
    const double *pdiag;
    bool transposed;  // A^T with (LU)^-T: the adjoint solve
    const double *tvals;  // values permuted into the transposed pattern (adjoint)
};

// Work space of one CG run: r, p, Ap, z (4n doubles) and two sets of reduction
// partials (2 ceil(n/256) doubles, sentinel-filled).  Null: the context's buffers.
struct cg_work {
    double *vec;
    double *part;
};

static void run_cg(hala_ctx c, const cg_problem &P, const double *b, double *x, double tol, int max_iter,
                   int *iterations, double *hist, int max_hist, const cg_work *w = nullptr) {
    const int n = P.A->nrows;
    HALA_REQUIRE(P.A->nrows == P.A->ncols, HALA_ERR_ARG, "CG needs a square matrix");
    HALA_REQUIRE(P.prec != PREC_ILU || (P.M && P.M->factored && P.M->n == n), HALA_ERR_ARG,
                 "CG: ILU preconditioner missing, unfactored or of the wrong size");
    HALA_REQUIRE(P.prec != PREC_JAC || P.pdiag, HALA_ERR_ARG, "CG: Jacobi preconditioner needs the diagonal");
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    // work vectors: r, p, Ap, z
    double *r, *part;
    if (w) {
        r = w->vec;
        part = w->part;
    } else {
        c->vec.reserve((size_t)4 * n * sizeof(double));
        r = c->vec.as<double>();
        part = partials64(c, (size_t)2 * nb);
    }
    double *p = r + n, *ap = p + n, *z = ap + n;
    double *part2 = part + nb;
    cg_state *st = cg_state_ptr(c);
    const int *pntr = P.transposed ? P.A->tp.tpntr.as<int>() : P.A->pntr;
    const int *indx = P.transposed ? P.A->tp.trow.as<int>() : P.A->indx;
    const double *vals = P.transposed ? P.tvals : P.vals;
    const short *boff = P.transposed ? nullptr : P.A->boff.as<short>();

    const int gs = (int)((n + SPMV_ROWS - 1) / SPMV_ROWS);
    // SpMV + dot: a workgroup per 256-row block, or grid-stride over them with the option's cap per CU
    // (128^3 PCG: 8 per CU 80.6 ms, 16 79.1 ms, a workgroup per block 78.7 ms)
    const int gsd = c->opt.spmv_blocks_per_cu > 0 ? std::min(gs, c->num_cus * c->opt.spmv_blocks_per_cu) : gs;
    const bool nt = stream_nt<double>(P.A->nnz);
    const int ge = (int)((nb + EK - 1) / EK);
    const int gp = grid_for(n, 256 * 4, 16384);
    cg_init_kernel<<<1, 1, 0, c->stream>>>(st, tol, max_iter, hist, max_hist);
    cg_residual_kernel<false><<<gs, 256, 0, c->stream>>>(n, pntr, indx, vals, x, b, r);
    HALA_CHECK_LAUNCH();
    const int *gate = &st->done;
    const bool fused = P.prec == PREC_ILU && !P.transposed && c->opt.cg_fuse && ilu_fused64(P.M);
    auto precond = [&](const double *in, double *out, const int *g) {
        if (P.transposed) ilu_apply_t64(P.M, in, out, g);
        else ilu_apply64(P.M, in, out, g);
    };
    switch (P.prec) {
    case PREC_ID: cg_start_kernel<PREC_ID><<<ge, 256, 0, c->stream>>>(n, r, nullptr, nullptr, p, part, wait_error_flag(c), st); break;
    case PREC_JAC:
        cg_start_kernel<PREC_JAC><<<ge, 256, 0, c->stream>>>(n, r, nullptr, P.pdiag, p, part, wait_error_flag(c), st);
        break;
    default:
        precond(r, z, nullptr);
        cg_start_kernel<PREC_ILU><<<ge, 256, 0, c->stream>>>(n, r, z, nullptr, p, part, wait_error_flag(c), st);
    }
    HALA_CHECK_LAUNCH();

    // small identity/Jacobi systems: the direction update rides in the next product (p and z's slot alternate)
    const bool pz = (P.prec == PREC_ID || P.prec == PREC_JAC) && !P.transposed && n <= PZ_MAX_ROWS;
    int pass = 0;
    auto one_pass = [&]() {
        if (pz) {
            double *pc = (pass & 1) ? z : p, *pn = (pass & 1) ? p : z;  // pass 0 uses p as cg_start left it
            if (pass == 0)
                cg_spmv_kernel<false><<<gs, 256, 0, c->stream>>>(n, pntr, indx, vals, pc, ap, part, wait_error_flag(c),
                                                                 st);
            else if (P.prec == PREC_ID)
                cg_spmv_pz_kernel<PREC_ID><<<gs, 256, 0, c->stream>>>(n, pntr, indx, vals, pn, r, nullptr, pc, ap,
                                                                      part, wait_error_flag(c), st);
            else
                cg_spmv_pz_kernel<PREC_JAC><<<gs, 256, 0, c->stream>>>(n, pntr, indx, vals, pn, r, P.pdiag, pc, ap,
                                                                       part, wait_error_flag(c), st);
            if (P.prec == PREC_ID)
                cg_update_kernel<PREC_ID><<<ge, 256, 0, c->stream>>>(n, pc, ap, nullptr, x, r, part, part2,
                                                                     wait_error_flag(c), st);
            else
                cg_update_kernel<PREC_JAC><<<ge, 256, 0, c->stream>>>(n, pc, ap, P.pdiag, x, r, part, part2,
                                                                      wait_error_flag(c), st);
            HALA_CHECK_LAUNCH();
            pass++;
            return;
        }
        if (boff) {  // columns as int16 offsets from the block's first row
            if (nt)
                cg_spmv_kernel<true, short><<<gsd, 256, 0, c->stream>>>(n, pntr, boff, vals, p, ap, part,
                                                                        wait_error_flag(c), st);
            else
                cg_spmv_kernel<false, short><<<gsd, 256, 0, c->stream>>>(n, pntr, boff, vals, p, ap, part,
                                                                         wait_error_flag(c), st);
        } else if (nt) {
            cg_spmv_kernel<true><<<gsd, 256, 0, c->stream>>>(n, pntr, indx, vals, p, ap, part, wait_error_flag(c), st);
        } else {
            cg_spmv_kernel<false><<<gsd, 256, 0, c->stream>>>(n, pntr, indx, vals, p, ap, part, wait_error_flag(c), st);
        }
        switch (P.prec) {
        case PREC_ID:
            cg_update_kernel<PREC_ID><<<ge, 256, 0, c->stream>>>(n, p, ap, nullptr, x, r, part, part2, wait_error_flag(c), st);
            cg_pupdate_kernel<PREC_ID><<<gp, 256, 0, c->stream>>>(n, nullptr, r, nullptr, p, st);
            break;
        case PREC_JAC:
            cg_update_kernel<PREC_JAC><<<ge, 256, 0, c->stream>>>(n, p, ap, P.pdiag, x, r, part, part2, wait_error_flag(c), st);
            cg_pupdate_kernel<PREC_JAC><<<gp, 256, 0, c->stream>>>(n, nullptr, r, P.pdiag, p, st);
            break;
        default:
            if (fused) {  // the sweeps' gather and scatter passes fused into the update and r.z
                tri_plan &L = P.M->L, &U = P.M->U;
                const int xu = c->opt.cg_xcd ? 1 : 0;
                cg_update_kernel<PREC_ILU><<<xu ? xcd_grid(ge) : ge, 256, 0, c->stream>>>(
                    n, p, ap, nullptr, nullptr, r, part, part2, wait_error_flag(c), st, L.slot_of_row.as<int>(),
                    L.col.xs64.as<double>(), xu);
                const col_axpy ax{x, p, &st->alpha, n};  // x += alpha p on the L sweep's helper workgroups
                ilu_apply64_mid(P.M, gate, &ax);
                cg_zscatter_rz_kernel<<<ge, 256, 0, c->stream>>>(n, U.slot_of_row.as<int>(), U.out64.as<double>(), r, z,
                                                                 part, wait_error_flag(c), st);
            } else {
                cg_update_kernel<PREC_ILU><<<ge, 256, 0, c->stream>>>(n, p, ap, nullptr, x, r, part, part2,
                                                                      wait_error_flag(c), st);
                precond(r, z, gate);
                cg_rz_kernel<<<ge, 256, 0, c->stream>>>(n, r, z, part, wait_error_flag(c), st);
            }
            cg_pupdate_kernel<PREC_ILU><<<gp, 256, 0, c->stream>>>(n, z, nullptr, nullptr, p, st);
        }
        HALA_CHECK_LAUNCH();
    };
    int batch = 4;
    int it = 1, done = 0, status = 0;
    double rn = 0.0;
    while (!done) {
        for (int k = 0; k < batch; k++) one_pass();
        cg_read(c, &it, &done, &status, &rn);
        if (batch < c->opt.cg_batch_max) batch = std::min(2 * batch, std::max(c->opt.cg_batch_max, 1));
    }
    if (fused) {  // the stopping pass's x update (its L sweep was gated off)
        cg_xfinal<<<gp, 256, 0, c->stream>>>(n, x, p, st);
        HALA_CHECK_LAUNCH();
    }
    check_wait_error(c);
    if (iterations) *iterations = it;
    HALA_REQUIRE(status == 0, HALA_ERR_NOCONV, "CG: residual norm became non-finite");
}

__global__ void sentinel_fill(double *__restrict__ x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = sentinel64();
}

// CG of hala_cg_f64 over a caller-provided work space of cg_noalloc_doubles(n)
// doubles: r, p, Ap, z, then the reduction partials.  The partials are re-armed
// (filled with the sentinel by a kernel) before the run; nothing is allocated.
void cg_noalloc(hala_csr A, const double *vals, int prec, hala_ilu ilu, const double *pdiag, const double *b,
                double *x, double tol, int max_iter, int *iterations, double *hist, int max_hist, double *work) {
    hala_ctx c = A->ctx;
    const int n = A->nrows;
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    cg_work w{work, work + (size_t)4 * n};
    if (n > 0) {
        sentinel_fill<<<grid_for(2 * nb, 256, 1024), 256, 0, c->stream>>>(w.part, 2 * nb);
        HALA_CHECK_LAUNCH();
    }
    cg_problem P{A, vals, prec, ilu, pdiag, false, nullptr};
    run_cg(c, P, b, x, tol, max_iter, iterations, hist, max_hist, &w);
}

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

int hala_cg_f64(hala_csr A, const double *vals, int prec, hala_ilu ilu, const double *pdiag, const double *b,
                double *x, double tol, int max_iter, int *iterations, double *hist, int max_hist) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A && vals && b && x, HALA_ERR_ARG, "hala_cg_f64: null argument");
        HALA_REQUIRE(prec >= 0 && prec <= 2, HALA_ERR_ARG, "hala_cg_f64: prec must be 0, 1 or 2");
        HALA_REQUIRE(!hist || max_hist >= 0, HALA_ERR_ARG, "hala_cg_f64: bad history size");
        set_device(A->ctx);
        cg_problem P{A, vals, prec, ilu, pdiag, false, nullptr};
        run_cg(A->ctx, P, b, x, tol, max_iter, iterations, hist, hist ? max_hist : 0);
    });
}

__global__ void permute_vals(int64_t nnz, const int *__restrict__ tmap, const double *__restrict__ v,
                             double *__restrict__ tv) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) tv[e] = v[tmap[e]];
}

int hala_cg_t_f64(hala_csr A, const double *vals, hala_ilu ilu, const double *b, double *x, double tol,
                  int max_iter, int *iterations, double *hist, int max_hist) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A && vals && ilu && b && x, HALA_ERR_ARG, "hala_cg_t_f64: null argument");
        HALA_REQUIRE(ilu->factored && ilu->n == A->nrows, HALA_ERR_ARG, "hala_cg_t_f64: ILU not factored / wrong size");
        hala_ctx c = A->ctx;
        set_device(c);
        ensure_transpose(A);
        ilu_prepare_adjoint(ilu);
        dbuf tv;
        tv.reserve((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(double));
        if (A->nnz > 0) {
            permute_vals<<<(int)((A->nnz + 255) / 256), 256, 0, c->stream>>>(A->nnz, A->tp.tmap.as<int>(), vals,
                                                                             tv.as<double>());
            HALA_CHECK_LAUNCH();
        }
        cg_problem P{A, vals, PREC_ILU, ilu, nullptr, true, tv.as<double>()};
        run_cg(c, P, b, x, tol, max_iter, iterations, hist, hist ? max_hist : 0);
        HALA_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
