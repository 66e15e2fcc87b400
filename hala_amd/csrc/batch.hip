// batch.hip -- thousands of independent small SPD systems solved by
// Jacobi-preconditioned CG in one persistent launch, replayed from a HIP graph.
//
// Each workgroup (a "team" of T threads) takes systems from a queue ordered
// largest first and runs the whole CG loop of one system with HALA's step
// order (hex/solvers/hala_solvers_cg.hpp:92-156; Jacobi apply = vdivide,
// wax/hala_blas_extensions.hpp:139-151).  Thread t owns rows t, t+T, ...; p
// lives in LDS (the product gathers it), r, x and A p stay in HBM-backed
// per-system slices.  Team dot products: thread t sums its rows in order from
// 0, then the team is reduced with the halving tree (pairs t, t+off).
// The team size and workgroups per CU come from hala_batch_tune().
#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <vector>

#include "common.hpp"

struct hala_batch_s {
    hala_ctx ctx = nullptr;
    int nsys = 0, n_total = 0, max_n = 0;
    int64_t nnz_total = 0;
    const int *row_off = nullptr, *pntr = nullptr, *indx = nullptr;  // device, not owned
    hala_amd::dbuf order, work, iters, ctl;
    int team = 256, wg_per_cu = 2;
    // graph cache
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *key[4] = {nullptr, nullptr, nullptr, nullptr};
    double key_rtol = -1.0;
    int key_max_iter = -1, key_team = -1, key_wg = -1;
    ~hala_batch_s() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

namespace hala_amd {

// halving-tree reduction over the team; every thread gets the result
template <int T>
__device__ __forceinline__ double team_sum(double v, double *red) {
    const int t = threadIdx.x;
    if (T > 64) {
        red[t] = v;
        __syncthreads();
#pragma unroll
        for (int off = T / 2; off >= 64; off >>= 1) {
            if (t < off) red[t] = red[t] + red[t + off];
            __syncthreads();
        }
        if (t < 64) v = red[t];
    }
    if (t < 64) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_down(v, off, 64);
        if (t == 0) red[0] = v;
    }
    __syncthreads();
    double r = red[0];
    __syncthreads();
    return r;
}

__global__ void batch_reset_kernel(unsigned *ctl) {
    ctl[0] = 0u;
    ctl[1] = 0u;
}

template <int T>
__global__ __launch_bounds__(T) void batch_pcg_kernel(int nsys, int n_total, int max_n, const int *__restrict__ order,
                                                      const int *__restrict__ row_off, const int *__restrict__ pntr,
                                                      const int *__restrict__ indx, const double *__restrict__ vals,
                                                      const double *__restrict__ dvals, const double *__restrict__ b,
                                                      double *__restrict__ x, double *__restrict__ work, double rtol,
                                                      int max_iter, int *__restrict__ iters, unsigned *ctl) {
    extern __shared__ double lds[];
    double *sp = lds;          // p of the current system
    double *red = lds + max_n; // team reduction scratch
    __shared__ int s_q;
    const int t = threadIdx.x;
    for (;;) {
        if (t == 0) s_q = (int)atomicAdd(ctl, 1u);
        __syncthreads();
        const int q = s_q;
        __syncthreads();
        if (q >= nsys) break;
        const int s = order[q];
        const int r0 = row_off[s];
        const int n = row_off[s + 1] - r0;
        const int *P = pntr + r0;
        double *xs = x + r0;
        double *rs = work + r0;
        double *aps = work + n_total + r0;
        const double *bs = b + r0;
        const double *ds = dvals + r0;
        // tol = rtol * norm2(b)
        double acc = 0.0;
        for (int i = t; i < n; i += T) acc = acc + bs[i] * bs[i];
        const double tol = rtol * sqrt(team_sum<T>(acc, red));
        // r = b + (-1)(1*(A x) + 0); z = r ./ d; p = z; zr = <r, z>
        acc = 0.0;
        for (int i = t; i < n; i += T) {
            double sum = 0.0;
            for (int j = P[i]; j < P[i + 1]; j++) sum = sum + vals[j] * xs[indx[j]];
            double ri = bs[i] + (-1.0) * (1.0 * sum + 0.0);
            rs[i] = ri;
            double zi = ri / ds[i];
            sp[i] = zi;
            acc = acc + ri * zi;
        }
        double zr = team_sum<T>(acc, red);  // (its barriers also publish sp)
        int it = 1;
        for (;;) {
            // Ap = A p, <p, Ap>
            acc = 0.0;
            for (int i = t; i < n; i += T) {
                double sum = 0.0;
                for (int j = P[i]; j < P[i + 1]; j++) sum = sum + vals[j] * sp[indx[j]];
                double y = 1.0 * sum + 0.0;
                aps[i] = y;
                acc = acc + sp[i] * y;
            }
            const double pap = team_sum<T>(acc, red);
            it++;
            const double a = zr / pap;
            double rr = 0.0, rz = 0.0;
            for (int i = t; i < n; i += T) {
                xs[i] = xs[i] + a * sp[i];
                double ri = rs[i] + (-a) * aps[i];
                rs[i] = ri;
                rr = rr + ri * ri;
                rz = rz + ri * (ri / ds[i]);
            }
            const double rn = sqrt(team_sum<T>(rr, red));
            const double nzr = team_sum<T>(rz, red);
            if (it == max_iter || rn < tol || !(rn == rn)) break;
            const double beta = nzr / zr;
            for (int i = t; i < n; i += T) sp[i] = beta * sp[i] + 1.0 * (rs[i] / ds[i]);
            zr = nzr;
            __syncthreads();  // p complete before the next product
        }
        if (t == 0) iters[s] = it;
        __syncthreads();
    }
}

static size_t batch_lds(int team, int max_n) { return (size_t)(max_n + team) * sizeof(double); }

static void launch_batch(hala_batch B, hipStream_t s, int team, int wg, const double *vals, const double *dvals,
                         const double *b, double *x, double rtol, int max_iter) {
    hala_ctx c = B->ctx;
    unsigned *ctl = B->ctl.as<unsigned>();
    int grid = std::max(1, std::min(B->nsys, c->num_cus * wg));
    size_t lds = batch_lds(team, B->max_n);
    batch_reset_kernel<<<1, 1, 0, s>>>(ctl);
#define HALA_BATCH(TT)                                                                                          \
    batch_pcg_kernel<TT><<<grid, TT, lds, s>>>(B->nsys, B->n_total, B->max_n, B->order.as<int>(), B->row_off,    \
                                               B->pntr, B->indx, vals, dvals, b, x, B->work.as<double>(), rtol,  \
                                               max_iter, B->iters.as<int>(), ctl)
    switch (team) {
    case 64: HALA_BATCH(64); break;
    case 128: HALA_BATCH(128); break;
    case 256: HALA_BATCH(256); break;
    case 512: HALA_BATCH(512); break;
    case 1024: HALA_BATCH(1024); break;
    default: throw status_error{HALA_ERR_ARG, "batch team size must be 64, 128, 256, 512 or 1024"};
    }
#undef HALA_BATCH
    HALA_CHECK_LAUNCH();
}

// one workgroup may use at most 64 KiB of dynamic LDS; a CU holds 160 KiB
static bool fits(hala_batch B, int team, int wg) {
    return batch_lds(team, B->max_n) <= 64 * 1024 && batch_lds(team, B->max_n) * wg <= 160 * 1024;
}

// Runs the batched solve through a captured graph (re-captured when the
// arguments or the launch shape change).
static void run_batch(hala_batch B, const double *vals, const double *dvals, const double *b, double *x, double rtol,
                      int max_iter) {
    hala_ctx c = B->ctx;
    bool same = B->exec && B->key[0] == vals && B->key[1] == dvals && B->key[2] == b && B->key[3] == x &&
                B->key_rtol == rtol && B->key_max_iter == max_iter && B->key_team == B->team &&
                B->key_wg == B->wg_per_cu;
    if (!same) {
        if (B->exec) (void)hipGraphExecDestroy(B->exec);
        if (B->graph) (void)hipGraphDestroy(B->graph);
        B->exec = nullptr;
        B->graph = nullptr;
        if (!B->cap) HALA_HIP(hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking));
        HALA_HIP(hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal));
        try {
            launch_batch(B, B->cap, B->team, B->wg_per_cu, vals, dvals, b, x, rtol, max_iter);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(B->cap, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HALA_HIP(hipStreamEndCapture(B->cap, &B->graph));
        HALA_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
        B->key[0] = vals;
        B->key[1] = dvals;
        B->key[2] = b;
        B->key[3] = x;
        B->key_rtol = rtol;
        B->key_max_iter = max_iter;
        B->key_team = B->team;
        B->key_wg = B->wg_per_cu;
    }
    HALA_HIP(hipGraphLaunch(B->exec, c->stream));
}

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

int hala_batch_create(hala_ctx c, int nsys, const int *row_off, const int *pntr, const int *indx, hala_batch *out) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(c && out && nsys > 0 && row_off && pntr && indx, HALA_ERR_ARG, "hala_batch_create: bad argument");
        *out = nullptr;
        set_device(c);
        std::vector<int> ro(nsys + 1);
        HALA_HIP(hipMemcpyAsync(ro.data(), row_off, (nsys + 1) * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HALA_HIP(hipStreamSynchronize(c->stream));
        HALA_REQUIRE(ro[0] == 0, HALA_ERR_ARG, "hala_batch_create: row_off[0] must be 0");
        int max_n = 0;
        for (int s = 0; s < nsys; s++) {
            HALA_REQUIRE(ro[s + 1] >= ro[s], HALA_ERR_ARG, "hala_batch_create: decreasing row offsets");
            max_n = std::max(max_n, ro[s + 1] - ro[s]);
        }
        HALA_REQUIRE((size_t)(max_n + 64) * sizeof(double) <= 64 * 1024, HALA_ERR_ARG,
                     "hala_batch_create: a system exceeds the LDS-resident size (8128 rows)");
        const int n_total = ro[nsys];
        // every system's pattern must stay inside its own rows
        int nnz_total = 0;
        HALA_HIP(hipMemcpyAsync(&nnz_total, pntr + n_total, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HALA_HIP(hipStreamSynchronize(c->stream));
        auto *B = new hala_batch_s();
        try {
            B->ctx = c;
            B->nsys = nsys;
            B->n_total = n_total;
            B->max_n = max_n;
            B->nnz_total = nnz_total;
            B->row_off = row_off;
            B->pntr = pntr;
            B->indx = indx;
            // largest systems first: the dynamic queue then balances the tail
            std::vector<int> ord(nsys);
            std::iota(ord.begin(), ord.end(), 0);
            std::stable_sort(ord.begin(), ord.end(),
                             [&](int a, int b) { return ro[a + 1] - ro[a] > ro[b + 1] - ro[b]; });
            B->order.reserve(nsys * sizeof(int));
            HALA_HIP(hipMemcpyAsync(B->order.p, ord.data(), nsys * sizeof(int), hipMemcpyHostToDevice, c->stream));
            B->work.reserve((size_t)2 * std::max(n_total, 1) * sizeof(double));
            B->iters.reserve(nsys * sizeof(int));
            B->ctl.reserve(4 * sizeof(unsigned));
            HALA_HIP(hipStreamSynchronize(c->stream));
            // kernel-side bounds: validate every system's columns
            std::vector<int> hp(n_total + 1), hi(nnz_total);
            HALA_HIP(hipMemcpy(hp.data(), pntr, (n_total + 1) * sizeof(int), hipMemcpyDeviceToHost));
            HALA_HIP(hipMemcpy(hi.data(), indx, (size_t)nnz_total * sizeof(int), hipMemcpyDeviceToHost));
            for (int s = 0; s < nsys; s++) {
                int ns = ro[s + 1] - ro[s];
                for (int i = ro[s]; i < ro[s + 1]; i++) {
                    HALA_REQUIRE(hp[i] <= hp[i + 1] && hp[i + 1] <= nnz_total, HALA_ERR_PATTERN,
                                 "hala_batch_create: bad row pointers at row " + std::to_string(i));
                    for (int j = hp[i]; j < hp[i + 1]; j++)
                        HALA_REQUIRE(hi[j] >= 0 && hi[j] < ns, HALA_ERR_PATTERN,
                                     "hala_batch_create: column outside its system at row " + std::to_string(i));
                }
            }
        } catch (...) {
            delete B;
            throw;
        }
        *out = B;
    });
}

int hala_batch_destroy(hala_batch B) {
    return guard([&] {
        if (!B) return;
        (void)hipSetDevice(B->ctx->device);
        (void)hipStreamSynchronize(B->ctx->stream);
        delete B;
    });
}

int hala_batch_set_shape(hala_batch B, int team, int wg_per_cu) {
    return guard([&] {
        HALA_REQUIRE(B, HALA_ERR_ARG, "null batch");
        HALA_REQUIRE(team == 64 || team == 128 || team == 256 || team == 512 || team == 1024, HALA_ERR_ARG,
                     "team must be 64..1024 (power of two)");
        HALA_REQUIRE(wg_per_cu >= 1 && wg_per_cu <= 16 && fits(B, team, wg_per_cu), HALA_ERR_ARG,
                     "workgroups per CU do not fit the LDS");
        B->team = team;
        B->wg_per_cu = wg_per_cu;
    });
}

int hala_batch_get_shape(hala_batch B, int *team, int *wg_per_cu) {
    return guard([&] {
        HALA_REQUIRE(B && team && wg_per_cu, HALA_ERR_ARG, "null argument");
        *team = B->team;
        *wg_per_cu = B->wg_per_cu;
    });
}

int hala_batch_pcg_f64(hala_batch B, const double *vals, const double *dvals, const double *b, double *x,
                       double rtol, int max_iter, int *iters) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(B && vals && dvals && b && x, HALA_ERR_ARG, "hala_batch_pcg_f64: null argument");
        HALA_REQUIRE(max_iter >= 2, HALA_ERR_ARG, "hala_batch_pcg_f64: max_iter must be >= 2");
        hala_ctx c = B->ctx;
        set_device(c);
        run_batch(B, vals, dvals, b, x, rtol, max_iter);
        if (iters)
            HALA_HIP(hipMemcpyAsync(iters, B->iters.p, B->nsys * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    });
}

// Autotuner: times every (team, workgroups per CU) shape that fits the LDS on
// the given problem (x reset to zero before each run, best of `reps`) and
// keeps the fastest.  ms_out (host, nullable) receives its time.
int hala_batch_tune(hala_batch B, const double *vals, const double *dvals, const double *b, double *x, double rtol,
                    int max_iter, int reps, int *team_out, int *wg_out, double *ms_out) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(B && vals && dvals && b && x && reps >= 1, HALA_ERR_ARG, "hala_batch_tune: bad argument");
        hala_ctx c = B->ctx;
        set_device(c);
        hipEvent_t e0, e1;
        HALA_HIP(hipEventCreate(&e0));
        HALA_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        int bt = B->team, bw = B->wg_per_cu;
        const int teams[] = {64, 128, 256, 512, 1024};
        const int wgs[] = {1, 2, 4, 8};
        try {
            for (int tm : teams)
                for (int wg : wgs) {
                    if (!fits(B, tm, wg) || tm * wg > 2048) continue;
                    float tbest = 1e30f;
                    for (int r = 0; r < reps; r++) {
                        HALA_HIP(hipMemsetAsync(x, 0, (size_t)B->n_total * sizeof(double), c->stream));
                        HALA_HIP(hipEventRecord(e0, c->stream));
                        launch_batch(B, c->stream, tm, wg, vals, dvals, b, x, rtol, max_iter);
                        HALA_HIP(hipEventRecord(e1, c->stream));
                        HALA_HIP(hipEventSynchronize(e1));
                        float ms = 0.f;
                        HALA_HIP(hipEventElapsedTime(&ms, e0, e1));
                        tbest = std::min(tbest, ms);
                    }
                    if (tbest < best) {
                        best = tbest;
                        bt = tm;
                        bw = wg;
                    }
                }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        B->team = bt;
        B->wg_per_cu = bw;
        if (team_out) *team_out = bt;
        if (wg_out) *wg_out = bw;
        if (ms_out) *ms_out = best;
    });
}

}  // extern "C"
