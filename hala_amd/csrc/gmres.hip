// gmres.hip -- restarted left-preconditioned GMRES resident on the device,
// and the FP32-inner / FP64-refinement solver built on it.
//
// Step order follows HALA's solve_gmres (hex/solvers/hala_solvers_gmres.hpp:
// 125-230): one classical Gram-Schmidt pass (W^T r, then r -= W c; the gemv
// pair of krylov_project, :67-72), norm and scaling, Givens rotations (drot /
// drotg), the packed upper Hessenberg, the absolute tolerance on |s_k Z_k|,
// the last rotation of a cycle not applied to Z (:210-216), and the
// triangular solve (dtpsv) plus x += W Z at the end of a cycle.  Here every
// scalar step runs on the device in the last block of the reduction that
// produces its input, so a restart cycle needs one host synchronisation.
#include <algorithm>
#include <cmath>
#include <vector>

#include "blas1.hpp"
#include "solvers.hpp"

namespace hala_amd {

constexpr int GK = 4;  // canonical chunks per block

struct gm_state {
    double tol, tol_rel, inner_res, outer_res, inv;
    double *hist;
    int k, nZ, nH, total, outer, max_outer, restart, in_cycle, all_done, nh, max_hist, first, status;
    int skip_inner;  // all_done || !in_cycle: gate of the inner-loop preconditioner
};

// device-side drotg / drot / dtpsv('U','N','N'), same formulas as the reference BLAS
__device__ void d_rotg(double *a, double *b, double *c, double *s) {
    double sa = *a, sb = *b;
    double roe = (fabs(sa) > fabs(sb)) ? sa : sb;
    double scale = fabs(sa) + fabs(sb);
    double r, z;
    if (scale == 0.0) {
        *c = 1.0;
        *s = 0.0;
        r = 0.0;
        z = 0.0;
    } else {
        double as = sa / scale, bs = sb / scale;
        r = scale * sqrt(as * as + bs * bs);
        r = (roe >= 0.0) ? r : -r;
        *c = sa / r;
        *s = sb / r;
        z = 1.0;
        if (fabs(sa) > fabs(sb)) z = *s;
        if (fabs(sb) >= fabs(sa) && *c != 0.0) z = 1.0 / *c;
    }
    *a = r;
    *b = z;
}

__device__ void d_rot(double *x, double *y, double c, double s) {
    double t = c * *x + s * *y;
    *y = c * *y - s * *x;
    *x = t;
}

__device__ void d_tpsv_un(int n, const double *ap, double *x) {
    int kk = n * (n + 1) / 2 - 1;
    for (int j = n - 1; j >= 0; j--) {
        if (x[j] != 0.0) {
            x[j] = x[j] / ap[kk];
            double t = x[j];
            int k = kk - 1;
            for (int i = j - 1; i >= 0; i--) {
                x[i] = x[i] - t * ap[k];
                k--;
            }
        }
        kk -= j + 1;
    }
}

struct gm_arrays {
    double *H, *S, *C, *Z, *cf;
};

__global__ void gm_init_kernel(gm_state *st, double tol, double tol_rel, int restart, int max_outer, double *hist,
                               int max_hist) {
    st->tol = tol;
    st->tol_rel = tol_rel;
    st->inner_res = 0.0;
    st->outer_res = tol + 1.0;
    st->inv = 0.0;
    st->hist = hist;
    st->k = st->nZ = st->nH = st->total = st->outer = 0;
    st->max_outer = max_outer;
    st->restart = restart;
    st->in_cycle = 0;
    st->skip_inner = 1;
    st->all_done = !(st->outer_res > tol && 0 < max_outer);
    st->nh = 0;
    st->max_hist = max_hist;
    st->first = 1;
    st->status = 0;
}

// t = b - A x with the reference's gemv('N', -1, x, 1, t) arithmetic (or t = b when x is known zero)
template <class TV>
__global__ __launch_bounds__(256) void gm_residual_kernel(int n, const int *__restrict__ pntr,
                                                          const int *__restrict__ indx, const TV *__restrict__ vals,
                                                          const double *__restrict__ x, const double *__restrict__ b,
                                                          double *__restrict__ t, const gm_state *st, int x_zero) {
    if (st->all_done) return;
    if (x_zero && st->outer == 0) {
        int i = blockIdx.x * SPMV_ROWS + threadIdx.x;
        if (i < n) t[i] = b[i];
        return;
    }
    __shared__ TV prod[spmv_tile<TV>::value];
    TV acc = csr_block_rowsum<TV, double, TV>(n, pntr, indx, vals, x, prod);
    int i = blockIdx.x * SPMV_ROWS + threadIdx.x;
    if (i >= n) return;
    double s = (sizeof(TV) == 4) ? (double)((TV)1 * acc + (TV)0) : (double)acc;
    t[i] = (-1.0) * s + 1.0 * b[i];
}

// t = A r (beta 0), gated on the inner loop (TI = short: the matrix's int16
// block offsets instead of its columns)
template <class TV, bool NT, class TI = int>
__global__ __launch_bounds__(256) void gm_spmv_kernel(int n, const int *__restrict__ pntr,
                                                      const TI *__restrict__ indx, const TV *__restrict__ vals,
                                                      const double *__restrict__ r, double *__restrict__ t,
                                                      const gm_state *st) {
    if (st->all_done || !st->in_cycle) return;
    __shared__ TV prod[spmv_tile<TV>::value];
    TV acc = csr_block_rowsum<TV, double, TV, false, NT, TI>(n, pntr, indx, vals, r, prod, blockIdx.x);
    int i = blockIdx.x * SPMV_ROWS + threadIdx.x;
    if (i >= n) return;
    double s = (sizeof(TV) == 4) ? (double)((TV)1 * acc + (TV)0) : (double)acc;
    t[i] = 1.0 * s + 0.0;
}

__global__ void gm_copy_kernel(int n, const double *__restrict__ src, double *__restrict__ dst, const gm_state *st,
                               int need_cycle) {
    if (st->all_done || (need_cycle && !st->in_cycle)) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

// Start of a cycle: inner_res = norm2(r); Z = [inner_res]; the inner loop runs
// while inner_res > tol and k < restart.
__global__ __launch_bounds__(256) void gm_cycle_norm_kernel(int n, const double *__restrict__ r, double *part,
                                                            int *err, gm_state *st, gm_arrays ga) {
    if (st->all_done) return;
    __shared__ double sh[CB * GK];
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    double v[GK];
#pragma unroll
    for (int q = 0; q < GK; q++) {
        int64_t i = ((int64_t)blockIdx.x * GK + q) * CB + threadIdx.x;
        v[q] = (i < n) ? r[i] * r[i] : 0.0;
    }
    tree256<double, GK>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < GK; q++) {
            int64_t b = (int64_t)blockIdx.x * GK + q;
            if (b < nb) st_agent(part + b, v[q]);
        }
    if (is_collector()) {
        double rr = final_pass<double>(part, nb, sh, err);
        if (threadIdx.x == 0) {
            double res = sqrt(rr);
            st->total = st->total + 1;
            if (st->first && st->tol_rel > 0.0) st->tol = st->tol_rel * res;
            st->first = 0;
            st->inner_res = res;
            st->inv = 1.0 / res;
            st->k = 0;
            st->nH = 0;
            ga.Z[0] = res;
            st->nZ = 1;
            st->in_cycle = (res > st->tol && 0 < st->restart) ? 1 : 0;
            st->skip_inner = !st->in_cycle;
        }
    }
}

// r = inv * r (dscal) and W[:, 0] = r
__global__ void gm_scale_first_kernel(int n, double *__restrict__ r, double *__restrict__ W, const gm_state *st) {
    if (st->all_done) return;
    const double a = st->inv;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = a * r[i];
        r[i] = v;
        W[i] = v;
    }
}

// cf[j] = cdot(W_j, r), one column per blockIdx.y (columns j <= k)
__global__ __launch_bounds__(256) void gm_mdot_kernel(int n, const double *__restrict__ W,
                                                      const double *__restrict__ r, double *part, int *err,
                                                      const gm_state *st, gm_arrays ga) {
    if (st->all_done || !st->in_cycle) return;
    const int col = blockIdx.y;
    if (col > st->k) return;
    __shared__ double sh[CB * GK];
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    const double *w = W + (int64_t)col * n;
    double v[GK];
#pragma unroll
    for (int q = 0; q < GK; q++) {
        int64_t i = ((int64_t)blockIdx.x * GK + q) * CB + threadIdx.x;
        v[q] = (i < n) ? w[i] * r[i] : 0.0;
    }
    tree256<double, GK>(v, sh);
    double *pc = part + (int64_t)col * nb;
    if (threadIdx.x == 0)
        for (int q = 0; q < GK; q++) {
            int64_t b = (int64_t)blockIdx.x * GK + q;
            if (b < nb) st_agent(pc + b, v[q]);
        }
    if (is_collector()) {
        double t = final_pass<double>(pc, nb, sh, err);
        if (threadIdx.x == 0) ga.cf[col] = 1.0 * t + 0.0;  // gemv('T', 1, W, r, 0, coeffs)
    }
}

// r = r - W cf (reference dgemv 'N' order: r_i += (-cf_j) W_ij, j ascending) fused
// with norm2(r); the last block then performs the host-side steps of the
// reference inner loop (rotations, new rotation, H column, residual estimate).
__global__ __launch_bounds__(256) void gm_orth_kernel(int n, const double *__restrict__ W, double *__restrict__ r,
                                                      double *part, int *err, gm_state *st, gm_arrays ga) {
    if (st->all_done || !st->in_cycle) return;
    __shared__ double sh[CB * GK];
    __shared__ double scf[GM_MAX_RESTART];
    const int k = st->k;
    for (int j = threadIdx.x; j <= k; j += blockDim.x) scf[j] = -ga.cf[j];
    __syncthreads();
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    double v[GK];
#pragma unroll
    for (int q = 0; q < GK; q++) {
        int64_t i = ((int64_t)blockIdx.x * GK + q) * CB + threadIdx.x;
        double rr = 0.0;
        if (i < n) {
            double yi = r[i];
            for (int j = 0; j <= k; j++) yi = yi + scf[j] * W[(int64_t)j * n + i];
            r[i] = yi;
            rr = yi * yi;
        }
        v[q] = rr;
    }
    tree256<double, GK>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < GK; q++) {
            int64_t b = (int64_t)blockIdx.x * GK + q;
            if (b < nb) st_agent(part + b, v[q]);
        }
    if (is_collector()) {
        double rr = final_pass<double>(part, nb, sh, err);
        if (threadIdx.x == 0) {
            double nrm = sqrt(rr);
            double *cf = ga.cf;
            st->total = st->total + 1;
            for (int i = 0; i < k; i++) d_rot(&cf[i], &cf[i + 1], ga.C[i], ga.S[i]);
            double c, s, beta = nrm;
            d_rotg(&cf[k], &beta, &c, &s);
            for (int j = 0; j <= k; j++) ga.H[st->nH + j] = cf[j];
            st->nH = st->nH + k + 1;
            ga.S[k] = s;
            ga.C[k] = c;
            double res = fabs(s * ga.Z[st->nZ - 1]);
            st->inner_res = res;
            if (st->hist && st->nh < st->max_hist) st->hist[st->nh] = res;
            st->nh = st->nh + 1;
            st->inv = 1.0 / nrm;
            int k1 = k + 1;
            st->k = k1;
            if (res > st->tol && k1 < st->restart) {
                ga.Z[st->nZ] = 0.0;
                st->nZ = st->nZ + 1;
                d_rot(&ga.Z[k1 - 1], &ga.Z[k1], c, s);
            } else {
                st->in_cycle = 0;
            }
            if (!isfinite(res)) {
                st->in_cycle = 0;
                st->status = 1;
            }
            st->skip_inner = !st->in_cycle;
        }
    }
}

// r = inv * r; W[:, k] = r while the cycle continues
__global__ void gm_scale_next_kernel(int n, double *__restrict__ r, double *__restrict__ W, const gm_state *st) {
    if (st->all_done || !st->in_cycle) return;
    const double a = st->inv;
    double *w = W + (int64_t)st->k * n;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = a * r[i];
        r[i] = v;
        w[i] = v;
    }
}

// end of a cycle: Z = H^-1 Z (packed upper), outer bookkeeping
__global__ void gm_cycle_end_kernel(gm_state *st, gm_arrays ga) {
    if (st->all_done) return;
    if (st->nH > 0) d_tpsv_un(st->nZ, ga.H, ga.Z);
    st->outer = st->outer + 1;
    st->outer_res = st->inner_res;
    if (st->status || !(st->outer_res > st->tol && st->outer < st->max_outer)) st->all_done = 2;  // 2: x update pending
}

// x += W Z (dgemv 'N' order) when the cycle produced a basis
__global__ void gm_combine_kernel(int n, const double *__restrict__ W, double *__restrict__ x, gm_state *st,
                                  const gm_arrays ga) {
    if (st->all_done == 1 || st->nH == 0) return;
    __shared__ double sz[GM_MAX_RESTART];
    const int nz = st->nZ;
    for (int j = threadIdx.x; j < nz; j += blockDim.x) sz[j] = 1.0 * ga.Z[j];
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double yi = x[i];
        for (int j = 0; j < nz; j++) yi = yi + sz[j] * W[(int64_t)j * n + i];
        x[i] = yi;
    }
}

// x = x + 1*d (axpy(1, d, x)) for the refinement update
__global__ void gm_axpy1_kernel(int n, const double *__restrict__ d, double *__restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = x[i] + 1.0 * d[i];
}

__global__ void gm_seal_kernel(gm_state *st) {
    if (st->all_done == 2) st->all_done = 1;
}

struct gm_problem {
    hala_csr A;
    const double *vals64;
    const float *vals32;  // non-null: FP32 operator (and FP32 ILU)
    hala_ilu M;           // null: identity preconditioner
};

// One GMRES solve; returns total preconditioner applications.
static void run_gmres(hala_ctx c, const gm_problem &P, const double *b, double *x, double tol, double tol_rel,
                      int x_zero, int restart, int max_outer, int *total, double *hist, int max_hist, int *nhist,
                      dbuf &work) {
    const int n = P.A->nrows;
    HALA_REQUIRE(P.A->nrows == P.A->ncols, HALA_ERR_ARG, "GMRES needs a square matrix");
    HALA_REQUIRE(restart >= 1 && restart <= GM_MAX_RESTART, HALA_ERR_ARG,
                 "GMRES restart must be in [1, " + std::to_string(GM_MAX_RESTART) + "]");
    HALA_REQUIRE(!P.M || (P.M->factored && P.M->n == n), HALA_ERR_ARG, "GMRES: ILU missing/unfactored/wrong size");
    HALA_REQUIRE(!P.vals32 || (P.M && P.M->has32), HALA_ERR_ARG, "GMRES FP32 inner solve needs an FP32 ILU factor");
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    // layout: t, r (n each), W (n*restart), state, H, S, C, Z, cf
    const size_t harr = (size_t)restart * (restart + 1) / 2 + 1;
    const size_t nd = (size_t)2 * n + (size_t)n * restart + 64 + harr + 4 * (size_t)(GM_MAX_RESTART + 1);
    work.reserve(nd * sizeof(double));
    double *t = work.as<double>(), *r = t + n, *W = r + n;
    gm_state *st = reinterpret_cast<gm_state *>(W + (size_t)n * restart);
    gm_arrays ga;
    ga.H = reinterpret_cast<double *>(st) + 64;
    ga.S = ga.H + harr;
    ga.C = ga.S + GM_MAX_RESTART + 1;
    ga.Z = ga.C + GM_MAX_RESTART + 1;
    ga.cf = ga.Z + GM_MAX_RESTART + 1;
    double *part = partials64(c, (size_t)(restart + 1) * nb);
    const int *pntr = P.A->pntr, *indx = P.A->indx;
    const short *boff = P.A->boff.as<short>();  // int16 block offsets of the columns (null: none)
    const int64_t nnz = P.A->nnz;
    const int gs = (int)((n + SPMV_ROWS - 1) / SPMV_ROWS);
    const int ge = (int)((nb + GK - 1) / GK);
    const int gp = grid_for(n, 256 * 4, 16384);
    gm_init_kernel<<<1, 1, 0, c->stream>>>(st, tol, tol_rel, restart, max_outer, hist, max_hist);
    HALA_CHECK_LAUNCH();
    int it_total = 0;
    int done = 0;
    while (!done) {
        // ---- start of a cycle
        if (P.vals32)
            gm_residual_kernel<float><<<gs, 256, 0, c->stream>>>(n, pntr, indx, P.vals32, x, b, t, st, x_zero);
        else
            gm_residual_kernel<double><<<gs, 256, 0, c->stream>>>(n, pntr, indx, P.vals64, x, b, t, st, x_zero);
        if (!P.M) gm_copy_kernel<<<gp, 256, 0, c->stream>>>(n, t, r, st, 0);
        else if (P.vals32) ilu_apply32x64(P.M, t, r, &st->all_done);
        else if (c->opt.ilu64_jacobi > 0) ilu_apply64_jacobi(P.M, t, r, &st->all_done, c->opt.ilu64_jacobi);
        else ilu_apply64(P.M, t, r, &st->all_done);
        gm_cycle_norm_kernel<<<ge, 256, 0, c->stream>>>(n, r, part, wait_error_flag(c), st, ga);
        gm_scale_first_kernel<<<gp, 256, 0, c->stream>>>(n, r, W, st);
        HALA_CHECK_LAUNCH();
        // ---- inner steps (gated on the device once the cycle stops)
        for (int s = 0; s < restart; s++) {
            if (P.vals32) {
                if (boff && stream_nt<float>(nnz))
                    gm_spmv_kernel<float, true, short><<<gs, 256, 0, c->stream>>>(n, pntr, boff, P.vals32, r, t, st);
                else if (boff)
                    gm_spmv_kernel<float, false, short><<<gs, 256, 0, c->stream>>>(n, pntr, boff, P.vals32, r, t, st);
                else if (stream_nt<float>(nnz)) gm_spmv_kernel<float, true><<<gs, 256, 0, c->stream>>>(n, pntr, indx, P.vals32, r, t, st);
                else gm_spmv_kernel<float, false><<<gs, 256, 0, c->stream>>>(n, pntr, indx, P.vals32, r, t, st);
            } else {
                if (boff && stream_nt<double>(nnz))
                    gm_spmv_kernel<double, true, short><<<gs, 256, 0, c->stream>>>(n, pntr, boff, P.vals64, r, t, st);
                else if (boff)
                    gm_spmv_kernel<double, false, short><<<gs, 256, 0, c->stream>>>(n, pntr, boff, P.vals64, r, t, st);
                else if (stream_nt<double>(nnz)) gm_spmv_kernel<double, true><<<gs, 256, 0, c->stream>>>(n, pntr, indx, P.vals64, r, t, st);
                else gm_spmv_kernel<double, false><<<gs, 256, 0, c->stream>>>(n, pntr, indx, P.vals64, r, t, st);
            }
            if (!P.M) gm_copy_kernel<<<gp, 256, 0, c->stream>>>(n, t, r, st, 1);
            else if (P.vals32) ilu_apply32x64(P.M, t, r, &st->skip_inner);
            else if (c->opt.ilu64_jacobi > 0) ilu_apply64_jacobi(P.M, t, r, &st->skip_inner, c->opt.ilu64_jacobi);
            else ilu_apply64(P.M, t, r, &st->skip_inner);
            gm_mdot_kernel<<<dim3(ge, restart), 256, 0, c->stream>>>(n, W, r, part, wait_error_flag(c), st,
                                                                      ga);
            gm_orth_kernel<<<ge, 256, 0, c->stream>>>(n, W, r, part + (size_t)restart * nb,
                                                      wait_error_flag(c), st, ga);
            gm_scale_next_kernel<<<gp, 256, 0, c->stream>>>(n, r, W, st);
            HALA_CHECK_LAUNCH();
        }
        // ---- end of the cycle
        gm_cycle_end_kernel<<<1, 1, 0, c->stream>>>(st, ga);
        gm_combine_kernel<<<gp, 256, 0, c->stream>>>(n, W, x, st, ga);
        gm_seal_kernel<<<1, 1, 0, c->stream>>>(st);
        HALA_CHECK_LAUNCH();
        gm_state h;
        HALA_HIP(hipMemcpyAsync(c->host_pinned, st, sizeof(gm_state), hipMemcpyDeviceToHost, c->stream));
        HALA_HIP(hipStreamSynchronize(c->stream));
        std::memcpy(&h, c->host_pinned, sizeof(gm_state));
        done = h.all_done;
        it_total = h.total;
        if (nhist) *nhist = std::min(h.nh, max_hist);
        if (done && h.status) {
            if (total) *total = it_total;
            throw status_error{HALA_ERR_NOCONV, "GMRES: residual estimate became non-finite"};
        }
    }
    check_wait_error(c);
    if (total) *total = it_total;
}

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

int hala_gmres_f64(hala_csr A, const double *vals, int prec, hala_ilu ilu, const double *b, double *x, double tol,
                   int restart, int max_outer, int *total, double *hist, int max_hist, int *nhist) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A && vals && b && x, HALA_ERR_ARG, "hala_gmres_f64: null argument");
        HALA_REQUIRE(prec == 0 || (prec == 1 && ilu), HALA_ERR_ARG, "hala_gmres_f64: prec 0 (identity) or 1 (ILU)");
        HALA_REQUIRE(!hist || max_hist >= 0, HALA_ERR_ARG, "hala_gmres_f64: bad history size");
        hala_ctx c = A->ctx;
        set_device(c);
        if (nhist) *nhist = 0;
        dbuf work;
        gm_problem P{A, vals, nullptr, prec == 1 ? ilu : nullptr};
        run_gmres(c, P, b, x, tol, 0.0, 0, restart, max_outer, total, hist, hist ? max_hist : 0, nhist, work);
    });
}

int hala_gmres_ir_f64(hala_csr A, const double *vals, const float *vals32, hala_ilu ilu, const double *b, double *x,
                      double rtol, double inner_rtol, int restart, int max_refine, int *total, double *hist,
                      int max_hist, int *nhist) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A && vals && ilu && b && x, HALA_ERR_ARG, "hala_gmres_ir_f64: null argument");
        HALA_REQUIRE(ilu->factored && (!vals32 || ilu->has32), HALA_ERR_ARG,
                     "hala_gmres_ir_f64: factor the ILU with want_f32 = 1 for the FP32 inner solve");
        HALA_REQUIRE(max_refine >= 0 && inner_rtol > 0.0 && inner_rtol < 1.0, HALA_ERR_ARG,
                     "hala_gmres_ir_f64: bad refinement parameters");
        hala_ctx c = A->ctx;
        set_device(c);
        const int n = A->nrows;
        const int inner_cycles = 20;
        dbuf vecs, work;
        vecs.reserve((size_t)2 * n * sizeof(double));
        double *r = vecs.as<double>(), *d = r + n;
        double *out = c->scal.as<double>() + S_DOT;
        auto norm_host = [&](const double *v) {
            cdot_launch<double>(c, n, v, v, out, 2);
            HALA_HIP(hipMemcpyAsync(c->host_pinned, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HALA_HIP(hipStreamSynchronize(c->stream));
            return c->host_pinned[0];
        };
        const double bn = norm_host(b);
        std::vector<double> hh;
        int tot = 0;
        for (int it = 0; it <= max_refine; it++) {
            HALA_HIP(hipMemcpyAsync(r, b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            spmv_axpby<double, double, double>(A, vals, x, -1.0, 1.0, r, false);
            const double rel = norm_host(r) / bn;
            hh.push_back(rel);
            if (rel < rtol || it == max_refine) break;
            HALA_HIP(hipMemsetAsync(d, 0, (size_t)n * sizeof(double), c->stream));
            gm_problem P{A, vals, vals32, ilu};
            int inner = 0;
            run_gmres(c, P, r, d, 0.0, inner_rtol, 1, restart, inner_cycles, &inner, nullptr, 0, nullptr, work);
            tot += inner;
            int g = grid_for(n, 256 * 4, 16384);
            hipLaunchKernelGGL(gm_axpy1_kernel, dim3(g), dim3(256), 0, c->stream, n, (const double *)d, x);
            HALA_CHECK_LAUNCH();
        }
        if (hist && max_hist > 0) {
            size_t m = std::min((size_t)max_hist, hh.size());
            HALA_HIP(hipMemcpyAsync(hist, hh.data(), m * sizeof(double), hipMemcpyHostToDevice, c->stream));
        }
        HALA_HIP(hipStreamSynchronize(c->stream));
        if (nhist) *nhist = (int)hh.size();
        if (total) *total = tot;
    });
}

}  // extern "C"
