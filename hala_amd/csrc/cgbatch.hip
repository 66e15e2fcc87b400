// cgbatch.hip -- multi-RHS preconditioned CG with one shared matrix, resident
// on the device: HALA's solve_batch_cg / solve_batch_cg_ilu
// (hex/solvers/hala_solvers_cg_batch.hpp:68-249), i.e. solve_cg_core with the
// batch operations of wax/hala_blas_extensions.hpp.  Column k of B, X is
// entries [k n, (k+1) n) (column-major, as the reference's sparse_gemm).  Every
// scalar is a vector over the columns:
//   r = b - A x; z = M r; p = z; zr_k = <r_k, z_k>; it = 1
//   loop: Ap = A p; it++; a_k = zr_k / <p_k, Ap_k>;
//         x_k += a_k p_k; r_k += (-a_k) Ap_k;
//         stop if it == max_iter or max_k ||r_k|| < tol (batch_max_norm2);
//         z = M r; nzr_k = <r_k, z_k>; b_k = nzr_k / zr_k; p_k = b_k p_k; p += 1*z; zr = nzr
// All columns iterate until the stop.  The dots are this build's canonical
// order per column; the product reads each matrix tile once for up to
// CGB_GROUP columns.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "blas1.hpp"
#include "solvers.hpp"

namespace hala_amd {

namespace {

constexpr int CGB_GROUP = 8;  // columns per product launch (row sums held in registers)
constexpr int EKB = 4;        // canonical 256-element chunks per block in the column kernels

struct cgb_state {
    double tol, rn;
    double *hist;
    int it, done, max_iter, max_hist;
};

__global__ void cgb_init(cgb_state *st, double tol, int max_iter, double *hist, int max_hist) {
    st->tol = tol;
    st->rn = 0.0;
    st->hist = hist;
    st->it = 1;
    st->done = 0;
    st->max_iter = max_iter;
    st->max_hist = max_hist;
}

// r_k = b_k + (-1)*(1*(A x_k) + 0)   (blockIdx.y = column)
__global__ __launch_bounds__(256) void cgb_residual(int n, const int *__restrict__ pntr, const int *__restrict__ indx,
                                                    const double *__restrict__ vals, const double *__restrict__ X,
                                                    const double *__restrict__ B, double *__restrict__ R) {
    __shared__ double prod[spmv_tile<double>::value];
    const int64_t o = (int64_t)blockIdx.y * n;
    double acc = csr_block_rowsum<double, double, double>(n, pntr, indx, vals, X + o, prod, blockIdx.x);
    int i = blockIdx.x * SPMV_ROWS + threadIdx.x;
    if (i < n) R[o + i] = B[o + i] + (-1.0) * (1.0 * acc + 0.0);
}

// Canonical partials of 256-element chunks: v[k] summed by tree256, stored to part[b].
template <int K>
__device__ __forceinline__ void store_partials(double (&v)[K], double *sh, double *part, int64_t nb) {
    tree256<double, K>(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < K; k++) {
            int64_t b = (int64_t)blockIdx.x * K + k;
            if (b < nb) st_agent(part + b, v[k]);
        }
}

// p_k = z_k (z = r for the identity), zr_k = <r_k, z_k>
template <bool IDENT>
__global__ __launch_bounds__(256) void cgb_start(int n, const double *__restrict__ R, const double *__restrict__ Z,
                                                 double *__restrict__ Pv, double *part, double *zr,
                                                 int *err) {
    __shared__ double sh[CB * EKB];
    const int64_t nb = ((int64_t)n + CB - 1) / CB, o = (int64_t)blockIdx.y * n;
    double v[EKB];
#pragma unroll
    for (int k = 0; k < EKB; k++) {
        int64_t i = ((int64_t)blockIdx.x * EKB + k) * CB + threadIdx.x;
        double a = 0.0;
        if (i < n) {
            double ri = R[o + i], zi = IDENT ? ri : Z[o + i];
            Pv[o + i] = zi;
            a = ri * zi;
        }
        v[k] = a;
    }
    double *pk = part + blockIdx.y * nb;
    store_partials<EKB>(v, sh, pk, nb);
    if (is_collector()) {
        double t = final_pass<double>(pk, nb, sh, err);
        if (threadIdx.x == 0) zr[blockIdx.y] = t;
    }
}

// Ap_k = A p_k for columns c0 .. c0+nc-1 (nc <= CGB_GROUP), each matrix tile
// read once; fused <p_k, Ap_k>, and block 0 forms a_k = zr_k / <p_k, Ap_k>.
__global__ __launch_bounds__(256) void cgb_spmm(int n, int c0, int nc, const int *__restrict__ pntr,
                                                const int *__restrict__ indx, const double *__restrict__ vals,
                                                const double *__restrict__ Pv, double *__restrict__ AP, double *part,
                                                const double *zr, double *al, cgb_state *st, int *err) {
    if (st->done) return;
    constexpr int TILE = spmv_tile<double>::value;
    __shared__ double prod[TILE];
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    const int r0 = blockIdx.x * SPMV_ROWS, row = r0 + threadIdx.x, rlast = min(r0 + SPMV_ROWS, n);
    const int bstart = pntr[r0], bend = pntr[rlast];
    int rs = 0, re = 0;
    if (row < n) {
        rs = pntr[row];
        re = pntr[row + 1];
    }
    double acc[CGB_GROUP];
#pragma unroll
    for (int k = 0; k < CGB_GROUP; k++) acc[k] = 0.0;
    for (int base = bstart; base < bend; base += TILE) {
        const int cnt = min(TILE, bend - base);
        const int lo = max(rs, base), hi = min(re, base + cnt);
#pragma unroll
        for (int k = 0; k < CGB_GROUP; k++) {
            if (k >= nc) break;
            const double *p = Pv + (int64_t)(c0 + k) * n;
            for (int e = threadIdx.x; e < cnt; e += SPMV_ROWS) prod[e] = vals[base + e] * p[indx[base + e]];
            __syncthreads();
            double a = acc[k];
            for (int j = lo; j < hi; j++) a = a + prod[j - base];
            acc[k] = a;
            __syncthreads();
        }
    }
    __shared__ double sh[CB];
#pragma unroll
    for (int k = 0; k < CGB_GROUP; k++) {
        if (k >= nc) break;
        const int64_t o = (int64_t)(c0 + k) * n;
        double v = 0.0;
        if (row < n) {
            const double y = 1.0 * acc[k] + 0.0;
            AP[o + row] = y;
            v = Pv[o + row] * y;
        }
        v = tree256_1<double>(v, sh);
        if (threadIdx.x == 0) st_agent(part + (int64_t)(c0 + k) * nb + blockIdx.x, v);
    }
    if (is_collector()) {
        for (int k = 0; k < nc; k++) {
            double t = final_pass<double>(part + (int64_t)(c0 + k) * nb, nb, sh, err);
            if (threadIdx.x == 0) al[c0 + k] = zr[c0 + k] / t;
        }
        if (threadIdx.x == 0 && c0 == 0) st->it = st->it + 1;
    }
}

// x_k += a_k p_k; r_k += (-a_k) Ap_k; rr_k = <r_k, r_k>
__global__ __launch_bounds__(256) void cgb_update(int n, const double *__restrict__ Pv, const double *__restrict__ AP,
                                                  double *__restrict__ X, double *__restrict__ R, double *part,
                                                  const double *al, double *rr, const cgb_state *st, int *err) {
    if (st->done) return;
    __shared__ double sh[CB * EKB];
    const int64_t nb = ((int64_t)n + CB - 1) / CB, o = (int64_t)blockIdx.y * n;
    const double a = al[blockIdx.y];
    double v[EKB];
#pragma unroll
    for (int k = 0; k < EKB; k++) {
        int64_t i = ((int64_t)blockIdx.x * EKB + k) * CB + threadIdx.x;
        double q = 0.0;
        if (i < n) {
            X[o + i] = X[o + i] + (1.0 * a) * Pv[o + i];
            double ri = R[o + i] + (-1.0 * a) * AP[o + i];
            R[o + i] = ri;
            q = ri * ri;
        }
        v[k] = q;
    }
    double *pk = part + blockIdx.y * nb;
    store_partials<EKB>(v, sh, pk, nb);
    if (is_collector()) {
        double t = final_pass<double>(pk, nb, sh, err);
        if (threadIdx.x == 0) rr[blockIdx.y] = t;
    }
}

// stop test on the largest <r_k, r_k> (batch_max_norm2: iamax, then its square root)
__global__ void cgb_stop(int nrhs, const double *rr, cgb_state *st) {
    if (st->done) return;
    double m = rr[0];
    for (int k = 1; k < nrhs; k++)
        if (rr[k] > m) m = rr[k];
    const double rn = sqrt(m);
    st->rn = rn;
    const int it = st->it;
    if (st->hist && it - 2 < st->max_hist) st->hist[it - 2] = rn;
    if (it == st->max_iter || rn < st->tol) st->done = 1;
}

// nzr_k = <r_k, z_k>; b_k = nzr_k / zr_k; zr_k = nzr_k
template <bool IDENT>
__global__ __launch_bounds__(256) void cgb_rz(int n, const double *__restrict__ R, const double *__restrict__ Z,
                                              double *part, double *zr, double *al, const cgb_state *st,
                                              int *err) {
    if (st->done) return;
    __shared__ double sh[CB * EKB];
    const int64_t nb = ((int64_t)n + CB - 1) / CB, o = (int64_t)blockIdx.y * n;
    double v[EKB];
#pragma unroll
    for (int k = 0; k < EKB; k++) {
        int64_t i = ((int64_t)blockIdx.x * EKB + k) * CB + threadIdx.x;
        double ri = i < n ? R[o + i] : 0.0;
        v[k] = (i < n) ? ri * (IDENT ? ri : Z[o + i]) : 0.0;
    }
    double *pk = part + blockIdx.y * nb;
    store_partials<EKB>(v, sh, pk, nb);
    if (is_collector()) {
        double t = final_pass<double>(pk, nb, sh, err);
        if (threadIdx.x == 0) {
            al[blockIdx.y] = t / zr[blockIdx.y];
            zr[blockIdx.y] = t;
        }
    }
}

// p_k = b_k p_k; p += 1*z   (batch_scal, then axpy over all columns)
template <bool IDENT>
__global__ void cgb_pupdate(int n, const double *__restrict__ R, const double *__restrict__ Z, double *__restrict__ Pv,
                            const double *al, const cgb_state *st) {
    if (st->done) return;
    const int64_t o = (int64_t)blockIdx.y * n;
    const double b = al[blockIdx.y];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double zi = IDENT ? R[o + i] : Z[o + i];
        Pv[o + i] = b * Pv[o + i] + 1.0 * zi;
    }
}

void cgb_read(hala_ctx c, cgb_state *st, int *it, int *done, double *rn) {
    HALA_HIP(hipMemcpyAsync(c->host_pinned, st, sizeof(cgb_state), hipMemcpyDeviceToHost, c->stream));
    HALA_HIP(hipStreamSynchronize(c->stream));
    cgb_state h;
    std::memcpy(&h, c->host_pinned, sizeof(cgb_state));
    *it = h.it;
    *done = h.done;
    if (rn) *rn = h.rn;
}

}  // namespace

void run_cg_batch(hala_ctx c, hala_csr A, const double *vals, hala_ilu M, int nrhs, const double *B, double *X,
                  double tol, int max_iter, int *iterations, double *hist, int max_hist) {
    const int n = A->nrows;
    HALA_REQUIRE(A->nrows == A->ncols, HALA_ERR_ARG, "batch CG needs a square matrix");
    HALA_REQUIRE(nrhs >= 1 && nrhs <= 65535, HALA_ERR_ARG, "batch CG: nrhs must lie in 1..65535");
    HALA_REQUIRE(!M || (M->factored && M->n == n), HALA_ERR_ARG, "batch CG: ILU missing, unfactored or of the wrong size");
    const bool ident = M == nullptr;
    const int64_t nb = ((int64_t)n + CB - 1) / CB;
    const size_t N = (size_t)n * nrhs;
    dbuf work, scal;
    work.reserve(4 * N * sizeof(double));  // R, P, AP, Z
    scal.reserve(sizeof(cgb_state) + 3 * (size_t)nrhs * sizeof(double));
    double *R = work.as<double>(), *Pv = R + N, *AP = Pv + N, *Z = AP + N;
    cgb_state *st = scal.as<cgb_state>();
    double *zr = reinterpret_cast<double *>(st + 1), *al = zr + nrhs, *rr = al + nrhs;
    double *part = partials64(c, (size_t)nrhs * nb);
    const int gs = (int)((n + SPMV_ROWS - 1) / SPMV_ROWS);
    const int ge = (int)((nb + EKB - 1) / EKB);
    const int gp = grid_for(n, 256 * 4, 4096);
    cgb_init<<<1, 1, 0, c->stream>>>(st, tol, max_iter, hist, max_hist);
    cgb_residual<<<dim3(gs, nrhs), 256, 0, c->stream>>>(n, A->pntr, A->indx, vals, X, B, R);
    HALA_CHECK_LAUNCH();
    const int *gate = &st->done;
    auto precond = [&](const int *g) {
        for (int k = 0; k < nrhs; k++) ilu_apply64(M, R + (size_t)k * n, Z + (size_t)k * n, g);
    };
    if (ident) {
        cgb_start<true><<<dim3(ge, nrhs), 256, 0, c->stream>>>(n, R, nullptr, Pv, part, zr, wait_error_flag(c));
    } else {
        precond(nullptr);
        cgb_start<false><<<dim3(ge, nrhs), 256, 0, c->stream>>>(n, R, Z, Pv, part, zr, wait_error_flag(c));
    }
    HALA_CHECK_LAUNCH();
    auto one_pass = [&]() {
        for (int c0 = 0; c0 < nrhs; c0 += CGB_GROUP)
            cgb_spmm<<<gs, 256, 0, c->stream>>>(n, c0, std::min(CGB_GROUP, nrhs - c0), A->pntr, A->indx, vals, Pv, AP,
                                                part, zr, al, st, wait_error_flag(c));
        cgb_update<<<dim3(ge, nrhs), 256, 0, c->stream>>>(n, Pv, AP, X, R, part, al, rr, st, wait_error_flag(c));
        cgb_stop<<<1, 1, 0, c->stream>>>(nrhs, rr, st);
        if (ident) {
            cgb_rz<true><<<dim3(ge, nrhs), 256, 0, c->stream>>>(n, R, nullptr, part, zr, al, st, wait_error_flag(c));
            cgb_pupdate<true><<<dim3(gp, nrhs), 256, 0, c->stream>>>(n, R, nullptr, Pv, al, st);
        } else {
            precond(gate);
            cgb_rz<false><<<dim3(ge, nrhs), 256, 0, c->stream>>>(n, R, Z, part, zr, al, st, wait_error_flag(c));
            cgb_pupdate<false><<<dim3(gp, nrhs), 256, 0, c->stream>>>(n, nullptr, Z, Pv, al, st);
        }
        HALA_CHECK_LAUNCH();
    };
    int batch = 4, it = 1, done = 0;
    while (!done) {
        for (int k = 0; k < batch; k++) one_pass();
        cgb_read(c, st, &it, &done, nullptr);
        if (batch < c->opt.cg_batch_max) batch = std::min(2 * batch, std::max(c->opt.cg_batch_max, 1));
    }
    check_wait_error(c);
    if (iterations) *iterations = it;
}

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

int hala_cg_batch_f64(hala_csr A, const double *vals, int prec, hala_ilu ilu, int nrhs, const double *B, double *X,
                      double tol, int max_iter, int *iterations, double *hist, int max_hist) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A && vals && B && X, HALA_ERR_ARG, "hala_cg_batch_f64: null argument");
        HALA_REQUIRE(prec == 0 || (prec == 1 && ilu), HALA_ERR_ARG,
                     "hala_cg_batch_f64: prec must be 0 (identity) or 1 (ILU, with its handle)");
        HALA_REQUIRE(!hist || max_hist >= 0, HALA_ERR_ARG, "hala_cg_batch_f64: bad history size");
        set_device(A->ctx);
        run_cg_batch(A->ctx, A, vals, prec == 1 ? ilu : nullptr, nrhs, B, X, tol, max_iter, iterations, hist,
                     hist ? max_hist : 0);
    });
}

}  // extern "C"
