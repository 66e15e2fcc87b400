// trexec.hip -- execution of the sync-free triangular sweeps (the persistent
// pool over level-ordered chunks).
#include <algorithm>
#include <climits>
#include <type_traits>

#include "trsolve.hpp"

namespace hala_amd {

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T wait_value(const T *p, int *err, int sleep) {
    T v = ld_agent(p);
    if (!is_sent(v)) return v;
    // exponential backoff between polls, capped at `sleep` slots of 64 cycles:
    // a lane far from ready polls rarely, which keeps the uncached traffic
    // (and so the latency seen by lanes about to become ready) low
    unsigned spins = 0;
    int nap = 1;
    for (;;) {
        for (int i = 0; i < nap; i++) __builtin_amdgcn_s_sleep(1);
        if (nap < sleep) nap = nap * 2 < sleep ? nap * 2 : sleep;
        v = ld_agent(p);
        if (!is_sent(v)) return v;
        ++spins;
        // another wave already gave up: stop waiting too
        if (spins > SPIN_LIMIT || ((spins & 1023u) == 0 && ld_agent(err) != 0)) {
            atomicOr(err, 1);
            return T(0);
        }
    }
}

// Everything one triangular-solve launch needs (passed by value).
template <class T, class TIN, class TOUT>
struct tri_args {
    int nchunk, nwaves, nlev, sleep;
    const int *chunk_row, *chunk_off, *chunk_len, *ent_dep;
    const T *ent_val, *dg_val;
    T alpha;
    const TIN *in;
    T *out;
    TOUT *out_nat;
    unsigned *ticket;  // nstream counters TK_STRIDE words apart; "waves finished" at TK_MAX * TK_STRIDE
    int nstream;
    int *err;
    const int *gate;
};

// Persistent pool over level-ordered chunks handed out by tickets: chunk
// k*S + s is the k-th ticket of stream s, and stream s is served by the waves
// with index s mod S inside their workgroup, so every resident workgroup
// serves every stream.  A chunk is only ever held by a running wave, and the
// lowest unfinished chunk always has every dependency done and a wave on it or
// about to take it (each stream hands its chunks out in increasing order), so
// no pool size deadlocks, also with the one-ahead ticket prefetch.  The last
// wave to leave re-arms the counters.  Per chunk a wave first loads everything
// that does not depend on other rows, then issues the loads of all its
// dependencies at once and only re-polls those still holding the sentinel.
template <class T, class TIN, class TOUT, int MODE, bool UNIT>
__global__ __launch_bounds__(256) void tri_exec_kernel(tri_args<T, TIN, TOUT> a) {
    if (a.gate && *a.gate) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int S = a.nstream;
    const int sid = (int)(threadIdx.x >> 6) % S;
    unsigned *tk = a.ticket + sid * TK_STRIDE;
    unsigned *left = a.ticket + TK_MAX * TK_STRIDE;
    int w = 0;
    if (lane == 0) w = (int)atomicAdd(tk, 1u) * S + sid;
    w = __shfl(w, 0, WAVE);
    while (w < a.nchunk) {
        int wn = 0;
        if (lane == 0) wn = (int)atomicAdd(tk, 1u) * S + sid;  // next ticket, latency overlapped
        const int slot = w * WAVE + lane;
        const int row = a.chunk_row[slot];
        const int len = a.chunk_len[w];
        const int64_t base = (int64_t)a.chunk_off[w] * WAVE + lane;
        // Everything a row needs besides its dependencies is loaded before the
        // wait, so that nothing but the dependency values sits between the
        // last producer's store and this row's own store.
        T xin = T(0), dgv = T(1);
        if (row >= 0) {
            xin = (T)a.in[row];
            if (!UNIT) dgv = a.dg_val[slot];
        }
        // TRI_SUB: acc = in, acc -= v*x term by term; TRI_SUM: acc = alpha*in - sum(v*x)
        auto publish = [&](T acc, T sum) {
            if (MODE == TRI_SUM) acc = a.alpha * xin - sum;
            if (!UNIT) acc = acc / dgv;
            st_agent(a.out + slot, acc);
            if (a.out_nat) a.out_nat[row] = (TOUT)acc;
        };
        if (len <= MAXD) {
            int d[MAXD];
            T v[MAXD], xv[MAXD];
#pragma unroll
            for (int k = 0; k < MAXD; k++) {
                d[k] = -1;
                if (row >= 0 && k < len) {
                    d[k] = a.ent_dep[base + (int64_t)k * WAVE];
                    v[k] = a.ent_val[base + (int64_t)k * WAVE];
                }
            }
            bool ready = true;
#pragma unroll
            for (int k = 0; k < MAXD; k++)
                if (d[k] >= 0) {
                    xv[k] = ld_agent(a.out + d[k]);
                    ready = ready && !is_sent(xv[k]);
                }
            // Each lane publishes its row as soon as its own dependencies are
            // in (not when the whole chunk is), and every dependency still
            // holding the sentinel is re-polled together, one round of loads
            // in flight per iteration.  The loop exit is wave-uniform.
            bool done = row < 0;
            unsigned spins = 0;
            for (;;) {
                if (!done && ready) {
                    T acc = xin, sum = T(0);
#pragma unroll
                    for (int k = 0; k < MAXD; k++)
                        if (d[k] >= 0) {
                            if (MODE == TRI_SUB) acc = acc - v[k] * xv[k];
                            else sum = sum + v[k] * xv[k];
                        }
                    publish(acc, sum);
                    done = true;
                }
                if (__all(done)) break;
                if (a.sleep > 0) __builtin_amdgcn_s_sleep(1);
                if (!done) {
                    ready = true;
#pragma unroll
                    for (int k = 0; k < MAXD; k++)
                        if (d[k] >= 0 && is_sent(xv[k])) {
                            xv[k] = ld_agent(a.out + d[k]);
                            ready = ready && !is_sent(xv[k]);
                        }
                }
                ++spins;
                // give up (flag the error, finish with zeros) after the spin
                // limit or as soon as another wave has given up
                bool quit = spins > SPIN_LIMIT;
                if ((spins & 1023u) == 0) quit = quit || ld_agent(a.err) != 0;
                if (quit && !done) {
                    atomicOr(a.err, 1);
#pragma unroll
                    for (int k = 0; k < MAXD; k++)
                        if (d[k] >= 0 && is_sent(xv[k])) xv[k] = T(0);
                    ready = true;
                }
            }
        } else if (row >= 0) {
            T acc = xin, sum = T(0);
            for (int k = 0; k < len; k++) {
                int dk = a.ent_dep[base + (int64_t)k * WAVE];
                if (dk < 0) break;
                T vk = a.ent_val[base + (int64_t)k * WAVE];
                T xk = wait_value(a.out + dk, a.err, a.sleep);
                if (MODE == TRI_SUB) acc = acc - vk * xk;
                else sum = sum + vk * xk;
            }
            publish(acc, sum);
        }
        w = __shfl(wn, 0, WAVE);
    }
    if (lane == 0) {
        unsigned dd = atomicAdd(left, 1u);
        if (dd == (unsigned)a.nwaves - 1) {
            if (a.nchunk > 0 && ld_agent(a.ticket) == 0u) atomicOr(a.err, 2);  // no wave ran the pool
            for (int k = 0; k < S; k++) st_agent(a.ticket + k * TK_STRIDE, 0u);
            st_agent(left, 0u);
        }
    }
}


template <class T, class TIN, class TOUT>
void tri_exec(hala_ctx c, tri_plan &p, int mode, T alpha, const TIN *in, TOUT *out_nat, const int *gate) {
    constexpr bool is64 = sizeof(T) == 8;
    HALA_REQUIRE(is64 ? p.has64 : p.has32, HALA_ERR_ARG, "triangular solve: values not loaded for this precision");
    static_assert(std::is_same<TIN, TOUT>::value, "input and output share a type");
    if (col_usable(c, p, is64)) {  // (its gather pass re-arms the polled output)
        tri_col_exec<T, TOUT>(c, p, mode, alpha, in, out_nat, gate);
        return;
    }
    tri_rearm_prec(c, p, is64);
    tri_args<T, TIN, TOUT> a;
    a.out = is64 ? reinterpret_cast<T *>(p.out64.p) : reinterpret_cast<T *>(p.out32.p);
    a.ent_val = is64 ? reinterpret_cast<const T *>(p.ent_val64.p) : reinterpret_cast<const T *>(p.ent_val32.p);
    a.dg_val = is64 ? reinterpret_cast<const T *>(p.dg_val64.p) : reinterpret_cast<const T *>(p.dg_val32.p);
    int pool = c->opt.tri_pool_waves > 0 ? c->opt.tri_pool_waves : c->num_cus * c->opt.tri_waves_per_cu;
    int blocks = (std::min(pool, (p.nchunk + 3) / 4 * 4) + 3) / 4;
    if (blocks < 1) blocks = 1;
    a.nchunk = p.nchunk;
    a.nlev = p.nlev;
    a.sleep = c->opt.tri_sleep;
    a.chunk_row = p.chunk_row.as<int>();
    a.chunk_off = p.chunk_off.as<int>();
    a.chunk_len = p.chunk_len.as<int>();
    a.ent_dep = p.ent_dep.as<int>();
    a.alpha = alpha;
    a.in = in;
    a.out_nat = out_nat;
    a.ticket = tri_tickets(c);
    a.err = wait_error_flag(c);
    a.gate = gate;
    using K = void (*)(tri_args<T, TIN, TOUT>);
    K kernel = mode == TRI_SUB ? (p.unit ? (K)tri_exec_kernel<T, TIN, TOUT, TRI_SUB, true>
                                         : (K)tri_exec_kernel<T, TIN, TOUT, TRI_SUB, false>)
                               : (p.unit ? (K)tri_exec_kernel<T, TIN, TOUT, TRI_SUM, true>
                                         : (K)tri_exec_kernel<T, TIN, TOUT, TRI_SUM, false>);
    a.nwaves = blocks * 4;
    a.nstream = std::max(1, std::min(c->opt.tri_streams, TK_MAX));
    kernel<<<blocks, 256, 0, c->stream>>>(a);
    HALA_CHECK_LAUNCH();
}

template void tri_exec<double, double, double>(hala_ctx, tri_plan &, int, double, const double *, double *, const int *);
template void tri_exec<float, double, double>(hala_ctx, tri_plan &, int, float, const double *, double *, const int *);
template void tri_exec<float, float, float>(hala_ctx, tri_plan &, int, float, const float *, float *, const int *);

}  // namespace hala_amd
