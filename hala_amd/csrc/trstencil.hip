// trstencil.hip -- pencil-scheduled ILU(0) sweeps of stencil matrices (see
// trstencil.hpp for the schedule).
//
// Row (i, j, k) = i + nx j + P k, P = nx ny.  Slots: the neighbour offsets
// (di, dj, dk) of a triangle, in CSR (ascending column) order; bit q of a
// row's mask says whether its slot-q entry exists.  L's row (i, j, k) runs at
// global step i + a j + b k, U's at the same step counted backwards; a slot's
// value was produced d = |di + a dj + b dk| >= 1 steps earlier.
//
// A workgroup is two waves working on one tile of TJ x TK lines:
//   wave 0 (compute): lane (jj, kk) owns line (j0 + jj, k0 + kk); per step it
//     forms its row: acc = b; acc = acc - c_q v_q for the present slots in
//     CSR order; U: acc = acc / d.  v_q comes from the lane's own ring of
//     recent results (di only), a neighbour lane's ring (other line of the
//     tile), or LDS (a line of another tile);
//   wave 1 (loader): polls the other tiles' results in global memory (value
//     as flag: a signalling-NaN sentinel means "not yet") some steps ahead and
//     stages them in an LDS ring, publishing its progress with a release store.
// Tiles take tickets in dependency order, so a tile's producers are always
// resident before it.  Every wait is bounded (error flag, then zeros).
//
// Sentinels: L's output y is all-sentinel when L starts (U re-arms each y[r]
// right after reading it), and L writes the sentinel into each row of U's
// output before U starts, so both sweeps poll their own outputs.
#include <algorithm>
#include <numeric>

#include "ilu.hpp"
#include "trstencil.hpp"
#include "util.hpp"

namespace hala_amd {

namespace {

// ---- stencil kinds: tile shape, level weights and lower slots (CSR order) ----
template <int KIND>
struct sk;
template <>
struct sk<SK_5> {  // 2-D 5-point
    static constexpr int TJ = 64, TK = 1, NS = 2, A = 1, B = 0;
    __host__ __device__ static constexpr int di(int q) { return q == 0 ? 0 : -1; }
    __host__ __device__ static constexpr int dj(int q) { return q == 0 ? -1 : 0; }
    __host__ __device__ static constexpr int dk(int) { return 0; }
};
template <>
struct sk<SK_9> {  // 2-D 9-point
    static constexpr int TJ = 64, TK = 1, NS = 4, A = 2, B = 0;
    __host__ __device__ static constexpr int di(int q) { return q == 3 ? -1 : q - 1; }
    __host__ __device__ static constexpr int dj(int q) { return q == 3 ? 0 : -1; }
    __host__ __device__ static constexpr int dk(int) { return 0; }
};
template <>
struct sk<SK_7> {  // 3-D 7-point
    static constexpr int TJ = 8, TK = 8, NS = 3, A = 1, B = 1;
    __host__ __device__ static constexpr int di(int q) { return q == 2 ? -1 : 0; }
    __host__ __device__ static constexpr int dj(int q) { return q == 1 ? -1 : 0; }
    __host__ __device__ static constexpr int dk(int q) { return q == 0 ? -1 : 0; }
};
template <>
struct sk<SK_27> {  // 3-D 27-point: 9 slots with dk = -1, 3 with (dk, dj) = (0, -1), then i-1
    static constexpr int TJ = 8, TK = 8, NS = 13, A = 2, B = 4;
    __host__ __device__ static constexpr int di(int q) { return q < 12 ? q % 3 - 1 : -1; }
    __host__ __device__ static constexpr int dj(int q) { return q < 9 ? q / 3 - 1 : (q < 12 ? -1 : 0); }
    __host__ __device__ static constexpr int dk(int q) { return q < 9 ? -1 : 0; }
};

// slot q of a triangle: lower slots as above; upper slots are the lower ones
// negated, in reverse order (ascending offset again)
template <int KIND, bool UP>
__host__ __device__ constexpr int sdi(int q) {
    return UP ? -sk<KIND>::di(sk<KIND>::NS - 1 - q) : sk<KIND>::di(q);
}
template <int KIND, bool UP>
__host__ __device__ constexpr int sdj(int q) {
    return UP ? -sk<KIND>::dj(sk<KIND>::NS - 1 - q) : sk<KIND>::dj(q);
}
template <int KIND, bool UP>
__host__ __device__ constexpr int sdk(int q) {
    return UP ? -sk<KIND>::dk(sk<KIND>::NS - 1 - q) : sk<KIND>::dk(q);
}
// steps back to the producer of slot q (>= 1)
template <int KIND, bool UP>
__host__ __device__ constexpr int sdelta(int q) {
    const int s = sdi<KIND, UP>(q) + sk<KIND>::A * sdj<KIND, UP>(q) + sk<KIND>::B * sdk<KIND, UP>(q);
    return UP ? s : -s;
}

// runtime view of a kind (host side and the packing kernel)
struct kind_info {
    int TJ, TK, NS, A, B;
    int di[13], dj[13], dk[13];
};
template <int KIND>
kind_info make_info() {
    kind_info k{sk<KIND>::TJ, sk<KIND>::TK, sk<KIND>::NS, sk<KIND>::A, sk<KIND>::B, {}, {}, {}};
    for (int q = 0; q < sk<KIND>::NS; q++) {
        k.di[q] = sk<KIND>::di(q);
        k.dj[q] = sk<KIND>::dj(q);
        k.dk[q] = sk<KIND>::dk(q);
    }
    return k;
}
kind_info info_of(int kind) {
    switch (kind) {
    case SK_5: return make_info<SK_5>();
    case SK_9: return make_info<SK_9>();
    case SK_7: return make_info<SK_7>();
    default: return make_info<SK_27>();
    }
}

// ---- detection ------------------------------------------------------------------
// With grid (nx, P = nx ny, nz = n / P): every entry of row r must be a
// neighbour of r in the 3 x 3 x 3 box, inside the grid.  st[0]: failure flag,
// st[1]: union of the neighbour bits (di+1) + 3 (dj+1) + 9 (dk+1).
__global__ void stencil_check(int n, int nx, int P, int nz, const int *__restrict__ pntr,
                              const int *__restrict__ indx, int *st) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int ny = P / nx;
    const int i = r % nx, j = (r / nx) % ny, k = r / P;
    unsigned bits = 0;
    bool bad = false;
    for (int e = pntr[r]; e < pntr[r + 1] && !bad; e++) {
        const long long off = (long long)indx[e] - r;
        int dk = 0;
        if (nz > 1) dk = off > P / 2 ? 1 : (off < -(P / 2) ? -1 : 0);
        const long long rem = off - (long long)dk * P;
        const int dj = rem > nx / 2 ? 1 : (rem < -(nx / 2) ? -1 : 0);
        const long long di = rem - (long long)dj * nx;
        if (di < -1 || di > 1) bad = true;
        else if (i + di < 0 || i + di >= nx || j + dj < 0 || j + dj >= ny || k + dk < 0 || k + dk >= nz) bad = true;
        else bits |= 1u << ((int)di + 1 + 3 * (dj + 1) + 9 * (dk + 1));
    }
    if (bad) atomicOr(st, 1);
    else atomicOr(st + 1, (int)bits);
}

// ---- packing ------------------------------------------------------------------
struct pack_args {
    int nx, ny, nz, P, A, B, TJ, NS, up;
    int di[13], dj[13], dk[13];
    const int4 *meta;
    const long long *base;
    const int *pntr, *indx, *diag;
    int *pos, *dpos;
    unsigned short *mask;
};

// one thread per (tile = blockIdx.y, step * 64 + lane): CSR positions of the row's slots
__global__ void stencil_pack(pack_args a) {
    const int tile = blockIdx.y;
    const int4 m = a.meta[tile];
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= m.w * 64) return;
    const int s = sl >> 6, lane = sl & 63;
    const int j = m.x + lane % a.TJ, k = m.y + lane / a.TJ;
    const int t = a.up ? m.z + m.w - 1 - s : m.z + s;
    const int i = t - a.A * j - a.B * k;
    const bool valid = j < a.ny && k < a.nz && i >= 0 && i < a.nx;
    const long long prow = a.base[tile] + sl;
    const int r = valid ? i + a.nx * j + a.P * k : -1;
    unsigned msk = 0;
    for (int q = 0; q < a.NS; q++) {
        int p = -1;
        if (valid) {
            const int ii = i + a.di[q], jn = j + a.dj[q], kn = k + a.dk[q];
            if (ii >= 0 && ii < a.nx && jn >= 0 && jn < a.ny && kn >= 0 && kn < a.nz) {
                const int col = ii + a.nx * jn + a.P * kn;
                for (int e = a.pntr[r]; e < a.pntr[r + 1]; e++)
                    if (a.indx[e] == col) p = e;
            }
        }
        if (p >= 0) msk |= 1u << q;
        a.pos[(a.base[tile] + (long long)s * 64) * a.NS + q * 64 + lane] = p;
    }
    a.mask[prow] = (unsigned short)msk;
    if (a.up) a.dpos[prow] = valid ? a.diag[r] : -1;
}

template <class V>
__global__ void stencil_values(int64_t count, const int *__restrict__ pos, const double *__restrict__ f,
                               V *__restrict__ out, V absent) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; e < count; e += (int64_t)gridDim.x * blockDim.x) {
        const int p = pos[e];
        out[e] = p >= 0 ? (V)f[p] : absent;
    }
}

template <class T>
__global__ void fill_sent(T *__restrict__ x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = sentinel<T>();
}

// ---- the sweep ---------------------------------------------------------------
template <class V, class TB, class TO, class TR>
struct sweep_args {
    int nx, ny, nz, P, ntiles;
    const int4 *meta;
    const long long *base;
    const int *tile_of_ticket;
    unsigned *ticket, *reset_ticket;
    const V *coef, *diag;
    const unsigned short *mask;
    const TB *bin;  // right-hand side, natural order
    TO *out;        // result, natural order (polled by the other tiles' loaders)
    TR *rearm;      // gets the sentinel at each row this sweep computes
    int *err;
    const int *gate;
};

constexpr unsigned ST_SPIN_LIMIT = 1u << 22;

template <class T>
__device__ __forceinline__ T st_poll(const T *p, int *err) {
    T v = ld_agent(p);
    unsigned spins = 0;
    while (is_sent(v)) {
        __builtin_amdgcn_s_sleep(1);
        v = ld_agent(p);
        ++spins;
        if (spins > ST_SPIN_LIMIT || ((spins & 1023u) == 0 && ld_agent(err) != 0)) {
            atomicOr(err, 1);
            return T(0);
        }
    }
    return v;
}

__device__ __forceinline__ double lane_shfl(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ float lane_shfl(float v, int src) { return __shfl(v, src, 64); }

template <int KIND, bool UP, class V, class TB, class TO, class TR>
__global__ __launch_bounds__(128) void stencil_sweep(sweep_args<V, TB, TO, TR> A) {
    using K = sk<KIND>;
    constexpr int NS = K::NS, TJ = K::TJ, TK = K::TK;
    constexpr int U = NS > 4 ? 8 : 16;   // unroll = ring length = prefetch distance (steps)
    constexpr int RL = NS > 4 ? 8 : 16;  // LDS ring of staged external values (steps)
    if (A.gate && *A.gate) return;
    __shared__ V ext[RL][NS][64];
    __shared__ int s_tile, s_ready, s_cons;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x == 0) {
        const unsigned tk = atomicAdd(A.ticket, 1u);
        if (tk < (unsigned)A.ntiles) {
            s_tile = A.tile_of_ticket[tk];
        } else {  // a counter that was not re-armed: flag it, run nothing
            s_tile = -1;
            atomicOr(A.err, 8);
        }
        s_ready = 0;
        s_cons = 0;
        if (tk == 0) *A.reset_ticket = 0u;  // the other triangle's counter, for its next sweep
    }
    __syncthreads();
    const int tile = s_tile;
    if (tile < 0) return;
    const int4 m = A.meta[tile];
    const long long base = A.base[tile];
    const int S = m.w;
    const int jj = lane % TJ, kk = lane / TJ;
    const int j = m.x + jj, k = m.y + kk;
    const bool active = j < A.ny && k < A.nz;
    const int tfirst = UP ? m.z + m.w - 1 : m.z;
    const int ibase = tfirst - K::A * j - K::B * k;  // i = ibase + s (L), ibase - s (U)
    const int rowbase = A.nx * j + A.P * k;
    // slot q's line lies in this tile (its value comes from a lane's ring)
    auto in_tile = [&](int q) {
        const int dj = sdj<KIND, UP>(q), dk = sdk<KIND, UP>(q);
        return jj + dj >= 0 && jj + dj < TJ && kk + dk >= 0 && kk + dk < TK;
    };
    auto slot_off = [&](int q) { return sdi<KIND, UP>(q) + A.nx * sdj<KIND, UP>(q) + A.P * sdk<KIND, UP>(q); };

    if (wave == 1) {
        // ---------------- loader: the other tiles' results, U steps ahead, staged in LDS
        V lq[U][NS];
        unsigned lm[U];
        auto issue = [&](int s, int u) {
            const int i = UP ? ibase - s : ibase + s;
            const bool valid = active && s < S && i >= 0 && i < A.nx;
            lm[u] = valid ? (unsigned)A.mask[base + (long long)s * 64 + lane] : 0u;
            const int r = i + rowbase;
#pragma unroll
            for (int q = 0; q < NS; q++) {
                lq[u][q] = V(0);
                if (valid && !in_tile(q) && ((lm[u] >> q) & 1u)) lq[u][q] = (V)ld_agent(A.out + (r + slot_off(q)));
            }
        };
#pragma unroll
        for (int u = 0; u < U; u++) issue(u, u);
        int cons = 0;
        for (int s0 = 0; s0 < S; s0 += U) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int s = s0 + u;
                if (s < S) {
                    // the LDS slot of step s is free once the compute wave has consumed step s - RL
                    for (unsigned spins = 0; cons < s - RL + 1; spins++) {
                        cons = __hip_atomic_load(&s_cons, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (cons >= s - RL + 1) break;
                        if (spins > ST_SPIN_LIMIT) {  // never expected: the compute wave stalled
                            atomicOr(A.err, 1);
                            cons = s;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    const int r = (UP ? ibase - s : ibase + s) + rowbase;
#pragma unroll
                    for (int q = 0; q < NS; q++) {
                        if (!in_tile(q) && ((lm[u] >> q) & 1u)) {
                            V v = lq[u][q];
                            if (is_sent(v)) v = (V)st_poll(A.out + (r + slot_off(q)), A.err);
                            ext[s % RL][q][lane] = v;
                        }
                    }
                    if (lane == 0)
                        __hip_atomic_store(&s_ready, s + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    issue(s + U, u);
                }
            }
        }
        return;
    }

    // ---------------- compute
    V ring[U];
    V qc[U][NS];
    V qd[U];
    TB qb[U];
    unsigned qm[U];
    auto fetch = [&](int s, int u) {
        const int i = UP ? ibase - s : ibase + s;
        const bool valid = active && s < S && i >= 0 && i < A.nx;
        if (s < S) {
            const long long pr = base + (long long)s * 64;
#pragma unroll
            for (int q = 0; q < NS; q++) qc[u][q] = __builtin_nontemporal_load(A.coef + pr * NS + q * 64 + lane);
            qm[u] = __builtin_nontemporal_load(A.mask + pr + lane);
            if (UP) qd[u] = __builtin_nontemporal_load(A.diag + pr + lane);
        }
        qb[u] = valid ? A.bin[i + rowbase] : TB(0);
    };
#pragma unroll
    for (int u = 0; u < U; u++) {
        ring[u] = V(0);
        qd[u] = V(1);
        fetch(u, u);
    }
    int rdy = 0;
    for (int s0 = 0; s0 < S; s0 += U) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int s = s0 + u;
            if (s < S) {
                for (unsigned spins = 0; rdy <= s; spins++) {
                    rdy = __hip_atomic_load(&s_ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (rdy > s) break;
                    if (spins > ST_SPIN_LIMIT) {  // never expected: the loader stalled
                        atomicOr(A.err, 1);
                        rdy = s + 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                const int i = UP ? ibase - s : ibase + s;
                const bool valid = active && i >= 0 && i < A.nx;
                V acc = (V)qb[u];
                const unsigned msk = qm[u];
#pragma unroll
                for (int q = 0; q < NS; q++) {
                    const int dj = sdj<KIND, UP>(q), dk = sdk<KIND, UP>(q);
                    const int d = sdelta<KIND, UP>(q);  // 1 <= d < U for every kind
                    const V own = ring[(u - d + U) % U];
                    V nv;
                    if (dj == 0 && dk == 0) {
                        nv = own;
                    } else {
                        const V sh = lane_shfl(own, lane + dj + TJ * dk);
                        nv = in_tile(q) ? sh : ext[s % RL][q][lane];
                    }
                    if ((msk >> q) & 1u) acc = acc - qc[u][q] * nv;
                }
                if (UP) acc = acc / qd[u];
                ring[u] = acc;
                if (valid) {
                    st_agent(A.out + (i + rowbase), (TO)acc);
                    A.rearm[i + rowbase] = sentinel<TR>();
                }
                if (lane == 0) __hip_atomic_store(&s_cons, s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                fetch(s + U, u);
            }
        }
    }
}

template <int KIND, class V, class TIN, class TOUT>
void sweep_pair(hala_ilu_s *M, const TIN *x, TOUT *r, const int *gate) {
    hala_ctx c = M->ctx;
    stencil_plan &sp = *M->sp;
    const bool f64 = sizeof(V) == 8;
    V *y = f64 ? reinterpret_cast<V *>(sp.y64.p) : reinterpret_cast<V *>(sp.y32.p);
    {
        stencil_tri &L = sp.L;
        sweep_args<V, TIN, V, TOUT> a{sp.nx, sp.ny, sp.nz, sp.P, L.ntiles, L.meta.as<int4>(), L.base.as<long long>(),
                                      L.tile_of_ticket.as<int>(), L.ticket.as<unsigned>(), sp.U.ticket.as<unsigned>(),
                                      reinterpret_cast<const V *>(f64 ? L.coef64.p : L.coef32.p), nullptr,
                                      L.mask.as<unsigned short>(), x, y, r, wait_error_flag(c), gate};
        stencil_sweep<KIND, false, V, TIN, V, TOUT><<<L.ntiles, 128, 0, c->stream>>>(a);
        HALA_CHECK_LAUNCH();
    }
    {
        stencil_tri &Up = sp.U;
        sweep_args<V, V, TOUT, V> a{sp.nx, sp.ny, sp.nz, sp.P, Up.ntiles, Up.meta.as<int4>(), Up.base.as<long long>(),
                                    Up.tile_of_ticket.as<int>(), Up.ticket.as<unsigned>(), sp.L.ticket.as<unsigned>(),
                                    reinterpret_cast<const V *>(f64 ? Up.coef64.p : Up.coef32.p),
                                    reinterpret_cast<const V *>(f64 ? Up.diag64.p : Up.diag32.p),
                                    Up.mask.as<unsigned short>(), y, r, y, wait_error_flag(c), gate};
        stencil_sweep<KIND, true, V, V, TOUT, V><<<Up.ntiles, 128, 0, c->stream>>>(a);
        HALA_CHECK_LAUNCH();
    }
}

// tiles, tickets and packed positions of one triangle
void build_tri(hala_ilu_s *M, stencil_tri &T, const kind_info &K, bool up) {
    hala_ctx c = M->ctx;
    stencil_plan &sp = *M->sp;
    const int ntiles = sp.ntj * sp.ntk;
    std::vector<int4> meta(ntiles);
    std::vector<long long> base(ntiles);
    std::vector<int> first(ntiles), last(ntiles);
    long long acc = 0;
    int maxS = 0;
    for (int tk = 0; tk < sp.ntk; tk++)
        for (int tj = 0; tj < sp.ntj; tj++) {
            const int t = tk * sp.ntj + tj;
            const int j0 = tj * K.TJ, k0 = tk * K.TK;
            const int j1 = std::min(j0 + K.TJ, sp.ny) - 1, k1 = std::min(k0 + K.TK, sp.nz) - 1;
            const int t0 = K.A * j0 + K.B * k0, t1 = (sp.nx - 1) + K.A * j1 + K.B * k1;
            meta[t] = make_int4(j0, k0, t0, t1 - t0 + 1);
            base[t] = acc;
            acc += (long long)(t1 - t0 + 1) * 64;
            first[t] = t0;
            last[t] = t1;
            maxS = std::max(maxS, t1 - t0 + 1);
        }
    // dependency order: L by first step, U by last step from the top
    std::vector<int> order(ntiles);
    std::iota(order.begin(), order.end(), 0);
    if (!up) std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return first[a] < first[b]; });
    else std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return last[a] > last[b]; });
    T.ntiles = ntiles;
    T.nprow = acc;
    T.meta.reserve((size_t)ntiles * sizeof(int4));
    T.base.reserve((size_t)ntiles * sizeof(long long));
    T.tile_of_ticket.reserve((size_t)ntiles * sizeof(int));
    T.ticket.reserve(sizeof(unsigned));
    HALA_HIP(hipMemcpyAsync(T.meta.p, meta.data(), meta.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    HALA_HIP(hipMemcpyAsync(T.base.p, base.data(), base.size() * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    HALA_HIP(hipMemcpyAsync(T.tile_of_ticket.p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice,
                            c->stream));
    HALA_HIP(hipMemsetAsync(T.ticket.p, 0, sizeof(unsigned), c->stream));
    T.pos.reserve((size_t)acc * K.NS * sizeof(int));
    T.mask.reserve((size_t)acc * sizeof(unsigned short));
    if (up) T.dpos.reserve((size_t)acc * sizeof(int));
    pack_args pa{};
    pa.nx = sp.nx;
    pa.ny = sp.ny;
    pa.nz = sp.nz;
    pa.P = sp.P;
    pa.A = K.A;
    pa.B = K.B;
    pa.TJ = K.TJ;
    pa.NS = K.NS;
    pa.up = up ? 1 : 0;
    for (int q = 0; q < K.NS; q++) {  // upper slots: the lower ones negated, in reverse order
        const int src = up ? K.NS - 1 - q : q;
        pa.di[q] = up ? -K.di[src] : K.di[src];
        pa.dj[q] = up ? -K.dj[src] : K.dj[src];
        pa.dk[q] = up ? -K.dk[src] : K.dk[src];
    }
    pa.meta = T.meta.as<int4>();
    pa.base = T.base.as<long long>();
    pa.pntr = M->pntr;
    pa.indx = M->indx;
    pa.diag = M->diag.as<int>();
    pa.pos = T.pos.as<int>();
    pa.dpos = up ? T.dpos.as<int>() : nullptr;
    pa.mask = T.mask.as<unsigned short>();
    dim3 g((unsigned)((maxS * 64 + 255) / 256), (unsigned)ntiles);
    stencil_pack<<<g, 256, 0, c->stream>>>(pa);
    HALA_CHECK_LAUNCH();
}

}  // namespace

bool stencil_build(hala_ilu_s *M) {
    hala_ctx c = M->ctx;
    const int n = M->n;
    if (n < 27) return false;
    // candidate grids from the column offsets of a few interior rows
    std::vector<std::pair<int, int>> cand;  // (nx, P)
    // (offsets by primes: rows at exact fractions of n tend to sit on a grid boundary)
    for (int64_t pr : {(int64_t)n / 2 + 1009, (int64_t)n / 3 + 2003, 2 * (int64_t)n / 3 + 3001, (int64_t)n / 4 + 577,
                       (int64_t)n / 2}) {
        const int probe = (int)(pr % (n - 1));
        int lo = fetch_i32(c, M->pntr + probe), hi = fetch_i32(c, M->pntr + probe + 1);
        if (hi - lo > 27 || hi <= lo) return false;
        std::vector<int> cols(hi - lo);
        HALA_HIP(hipMemcpyAsync(cols.data(), M->indx + lo, cols.size() * sizeof(int), hipMemcpyDeviceToHost,
                                c->stream));
        HALA_HIP(hipStreamSynchronize(c->stream));
        std::vector<long long> mag;
        for (int col : cols)
            if (col < probe) mag.push_back((long long)probe - col);
        std::sort(mag.begin(), mag.end());
        if (mag.empty()) continue;
        std::vector<int> nxs;
        for (long long u : mag)
            if (u >= 2) {
                nxs.push_back((int)u);
                nxs.push_back((int)u + 1);
                break;
            }
        for (int nx : nxs) {
            if (nx < 3 || n % nx) continue;
            std::vector<long long> Ps = {n};  // 2-D: one plane
            const long long umax = mag.back();
            for (long long d : {0LL, (long long)nx - 1, (long long)nx, (long long)nx + 1})
                if (umax - d > 2LL * nx + 2) Ps.push_back(umax - d);
            for (long long P : Ps)
                if (P % nx == 0 && n % P == 0 && P <= n) cand.push_back({nx, (int)P});
        }
    }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    if (cand.empty() || cand.size() > 24) return false;
    int *st = reinterpret_cast<int *>(c->scal.as<double>() + S_DOT + 8);
    for (auto [nx, P] : cand) {
        const int nz = n / P, ny = P / nx;
        if (nz > 1 && ny < 3) continue;
        HALA_HIP(hipMemsetAsync(st, 0, 2 * sizeof(int), c->stream));
        stencil_check<<<(n + 255) / 256, 256, 0, c->stream>>>(n, nx, P, nz, M->pntr, M->indx, st);
        HALA_CHECK_LAUNCH();
        int h[2];
        HALA_HIP(hipMemcpyAsync(c->host_pinned, st, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HALA_HIP(hipStreamSynchronize(c->stream));
        std::memcpy(h, c->host_pinned, sizeof(h));
        if (h[0]) continue;
        const unsigned bits = (unsigned)h[1];
        // the smallest kind holding every neighbour in use
        const auto subset = [&](std::initializer_list<int> s) {
            unsigned k = 1u << 13;
            for (int b : s) k |= 1u << b;
            return (bits & ~k) == 0;
        };
        int kind;
        if (nz == 1) {
            if (subset({10, 12, 14, 16})) kind = SK_5;
            else if (subset({9, 10, 11, 12, 14, 15, 16, 17})) kind = SK_9;
            else continue;
        } else {
            kind = subset({4, 10, 12, 14, 16, 22}) ? SK_7 : SK_27;
        }
        auto sp = std::make_unique<stencil_plan>();
        sp->kind = kind;
        sp->nx = nx;
        sp->ny = ny;
        sp->nz = nz;
        sp->P = P;
        const kind_info K = info_of(kind);
        sp->ntj = (ny + K.TJ - 1) / K.TJ;
        sp->ntk = (nz + K.TK - 1) / K.TK;
        M->sp = std::move(sp);
        build_tri(M, M->sp->L, K, false);
        build_tri(M, M->sp->U, K, true);
        M->sp->y64.reserve((size_t)n * sizeof(double));
        fill_sent<double><<<grid_for(n, 1024, 4096), 256, 0, c->stream>>>(M->sp->y64.as<double>(), n);
        HALA_CHECK_LAUNCH();
        return true;
    }
    return false;
}

void stencil_load(hala_ilu_s *M, bool f32) {
    hala_ctx c = M->ctx;
    stencil_plan &sp = *M->sp;
    const kind_info K = info_of(sp.kind);
    const double *f = M->factor.as<double>();
    auto load = [&](stencil_tri &T, bool up) {
        const int64_t nc = T.nprow * K.NS;
        T.coef64.reserve((size_t)nc * sizeof(double));
        stencil_values<double><<<grid_for(nc, 1024, 8192), 256, 0, c->stream>>>(nc, T.pos.as<int>(), f,
                                                                                T.coef64.as<double>(), 0.0);
        if (up) {
            T.diag64.reserve((size_t)T.nprow * sizeof(double));
            stencil_values<double><<<grid_for(T.nprow, 1024, 8192), 256, 0, c->stream>>>(
                T.nprow, T.dpos.as<int>(), f, T.diag64.as<double>(), 1.0);
        }
        if (f32) {
            T.coef32.reserve((size_t)nc * sizeof(float));
            stencil_values<float><<<grid_for(nc, 1024, 8192), 256, 0, c->stream>>>(nc, T.pos.as<int>(), f,
                                                                                  T.coef32.as<float>(), 0.0f);
            if (up) {
                T.diag32.reserve((size_t)T.nprow * sizeof(float));
                stencil_values<float><<<grid_for(T.nprow, 1024, 8192), 256, 0, c->stream>>>(
                    T.nprow, T.dpos.as<int>(), f, T.diag32.as<float>(), 1.0f);
            }
        }
        HALA_CHECK_LAUNCH();
    };
    load(sp.L, false);
    load(sp.U, true);
    if (f32 && !sp.y32.p) {
        sp.y32.reserve((size_t)M->n * sizeof(float));
        fill_sent<float><<<grid_for(M->n, 1024, 4096), 256, 0, c->stream>>>(sp.y32.as<float>(), M->n);
        HALA_CHECK_LAUNCH();
    }
    sp.loaded64 = true;
    sp.loaded32 = f32;
}

template <class V, class TIN, class TOUT>
void stencil_apply(hala_ilu_s *M, const TIN *x, TOUT *r, const int *gate) {
    switch (M->sp->kind) {
    case SK_5: sweep_pair<SK_5, V, TIN, TOUT>(M, x, r, gate); break;
    case SK_9: sweep_pair<SK_9, V, TIN, TOUT>(M, x, r, gate); break;
    case SK_7: sweep_pair<SK_7, V, TIN, TOUT>(M, x, r, gate); break;
    default: sweep_pair<SK_27, V, TIN, TOUT>(M, x, r, gate); break;
    }
}
template void stencil_apply<double, double, double>(hala_ilu_s *, const double *, double *, const int *);
template void stencil_apply<float, double, double>(hala_ilu_s *, const double *, double *, const int *);
template void stencil_apply<float, float, float>(hala_ilu_s *, const float *, float *, const int *);

}  // namespace hala_amd
