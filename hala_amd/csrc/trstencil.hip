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
    int *pos, *dpos, *prow;
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
    else a.prow[prow] = r;  // the packed vectors' row map (the L order)
}

// One thread per (step record g, lane): the factor's values of packed row
// g * 64 + lane into the record (coefficients, diagonal for U, mask).
template <class V>
__global__ void stencil_records(int64_t nrec, int NS, int up, int rec_bytes, const int *__restrict__ pos,
                                const int *__restrict__ dpos, const unsigned short *__restrict__ mask,
                                const double *__restrict__ f, unsigned char *__restrict__ rec) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < nrec * 64; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = t >> 6;
        const int lane = (int)(t & 63);
        unsigned char *r = rec + g * rec_bytes;
        V *coef = reinterpret_cast<V *>(r);
        for (int q = 0; q < NS; q++) {
            const int p = pos[(g * NS + q) * 64 + lane];
            coef[q * 64 + lane] = p >= 0 ? (V)f[p] : V(0);
        }
        int off = NS * 64 * (int)sizeof(V);
        if (up) {
            const int p = dpos[t];
            reinterpret_cast<V *>(r + off)[lane] = p >= 0 ? (V)f[p] : V(1);
            off += 64 * (int)sizeof(V);
        }
        reinterpret_cast<unsigned short *>(r + off)[lane] = mask[t];
    }
}

template <class T>
__global__ void fill_sent(T *__restrict__ x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = sentinel<T>();
}

// ---- the sweep ---------------------------------------------------------------
// Step record of a tile (written by stencil_records): the coefficients
// [NS][64], the diagonal [64] (U only) and the presence masks [64] (u16), so
// that the streaming wave moves a step's operands with NCH 16-byte loads per lane.
template <int KIND, bool UP, class V>
struct shape {
    static constexpr int NS = sk<KIND>::NS;
    static constexpr int OFF_DIAG = NS * 64 * (int)sizeof(V);
    static constexpr int OFF_MASK = OFF_DIAG + (UP ? 64 * (int)sizeof(V) : 0);
    static constexpr int REC = OFF_MASK + 128;       // bytes per step record (a multiple of 128)
    static constexpr int NCH = (REC + 1023) / 1024;  // 16-byte loads per lane and step
    // streamed steps in flight: (NCH + 1) loads each, at most 56 of a wave's 63
    static constexpr int U = 56 / (NCH + 1) < 16 ? 56 / (NCH + 1) : 16;
    static constexpr int UE = 8;                     // polled steps in flight
    static constexpr int RL = NCH > 4 ? 8 : 16;      // staged steps in LDS
    static constexpr int RG = 8;                     // compute unroll = own-result ring
    static constexpr int R = 32;                     // halo ring (steps), a power of two
    static constexpr int RS = R + 1;                 // halo line stride (odd: lanes of different lines hit different banks)
};

// Lines around a tile whose values its rows read ("frame"): in 3-D the 36
// lines (rj, rk) with rj in [-1, TJ], rk in [-1, TK] outside the tile, in 2-D
// the two lines beside it.
template <int KIND>
__host__ __device__ constexpr int frame_count() {
    return sk<KIND>::TK == 1 ? 2 : 2 * (sk<KIND>::TJ + 2) + 2 * sk<KIND>::TK;
}
template <int KIND>
__host__ __device__ constexpr int frame_id(int rj, int rk) {
    constexpr int TJ = sk<KIND>::TJ, TK = sk<KIND>::TK, W = TJ + 2;
    if (TK == 1) return rk != 0 ? -1 : (rj == -1 ? 0 : (rj == TJ ? 1 : -1));
    if (rk == -1 && rj >= -1 && rj <= TJ) return rj + 1;
    if (rk == TK && rj >= -1 && rj <= TJ) return W + rj + 1;
    if (rj == -1 && rk >= 0 && rk < TK) return 2 * W + rk;
    if (rj == TJ && rk >= 0 && rk < TK) return 2 * W + TK + rk;
    return -1;
}
template <int KIND>
__host__ __device__ constexpr int frame_rj(int f) {
    constexpr int TJ = sk<KIND>::TJ, TK = sk<KIND>::TK, W = TJ + 2;
    if (TK == 1) return f == 0 ? -1 : TJ;
    return f < W ? f - 1 : (f < 2 * W ? f - W - 1 : (f < 2 * W + TK ? -1 : TJ));
}
template <int KIND>
__host__ __device__ constexpr int frame_rk(int f) {
    constexpr int TJ = sk<KIND>::TJ, TK = sk<KIND>::TK, W = TJ + 2;
    if (TK == 1) return 0;
    return f < W ? -1 : (f < 2 * W ? TK : (f < 2 * W + TK ? f - 2 * W : f - 2 * W - TK));
}
// bit f: some lane of the tile reads frame line f through some slot
template <int KIND, bool UP>
constexpr unsigned long long frame_used() {
    using K = sk<KIND>;
    unsigned long long m = 0;
    for (int kk = 0; kk < K::TK; kk++)
        for (int jj = 0; jj < K::TJ; jj++)
            for (int q = 0; q < K::NS; q++) {
                const int rj = jj + sdj<KIND, UP>(q), rk = kk + sdk<KIND, UP>(q);
                if (rj < 0 || rj >= K::TJ || rk < 0 || rk >= K::TK) m |= 1ull << frame_id<KIND>(rj, rk);
            }
    return m;
}
template <int KIND, bool UP>
constexpr int delta_max() {
    int d = 0;
    for (int q = 0; q < sk<KIND>::NS; q++) d = sdelta<KIND, UP>(q) > d ? sdelta<KIND, UP>(q) : d;
    return d;
}

// Vectors of a sweep live in the packed order of the L records: row (i, j, k)
// of tile t, lane l, at L step s = i + A j + B k - first(t) sits at
// base(t) + 64 s + l, so a tile step reads and writes 512 contiguous bytes.
template <class V>
struct sweep_args {
    int nx, ny, nz, P, ntiles, ntj;
    long long nprow;
    const int4 *meta;
    const long long *base;
    const int *tile_of_ticket;
    unsigned *ticket, *reset_ticket;
    const unsigned char *rec;  // the triangle's step records in precision V
    const V *bin;              // right-hand side (packed)
    V *out;                    // result (packed; polled by the other tiles)
    V *rearm;                  // gets the sentinel at each packed row this sweep computes
    int *err;
    const int *gate;
};

constexpr unsigned ST_SPIN_LIMIT = 1u << 22;

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Polls a result until its producer has written it.  Loads only: on the spin
// limit, or when another wave has raised the error flag, sets bad and returns
// 0.  Every load it issues is examined before it returns, so no load of the
// poll is left in flight for the caller's waits to count.
template <class T>
__device__ __forceinline__ T st_poll(const T *p, const int *err, bool &bad) {
    T v = ld_agent(p);
    unsigned spins = 0;
    while (is_sent(v)) {
        if (++spins > ST_SPIN_LIMIT || ((spins & 1023u) == 0 && ld_agent(err) != 0)) {
            bad = true;
            v = T(0);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = ld_agent(p);
    }
    return v;
}

// Bounded LDS wait of one wave for a progress counter (another wave of the
// workgroup publishes it); false when the spin limit is reached.
__device__ __forceinline__ bool lds_wait(int *flag, int &seen, int want) {
    for (unsigned spins = 0; seen < want; spins++) {
        seen = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (seen >= want) break;
        if (spins > ST_SPIN_LIMIT) {
            seen = want;
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// Three waves per tile, each with one kind of memory traffic, so that the
// compiler's wait for a load counts only younger loads of the same stream:
//   wave 0 (compute): lane (jj, kk) owns line (j0 + jj, k0 + kk); per step it
//     forms its row from LDS: acc = b - sum of c_q v_q over the present slots
//     in CSR order (U: / d), v_q from its own register ring (i-1 terms) or the
//     halo ring; stores the result and re-arms the row of the other vector;
//   wave 1 (streamer): the step records and right-hand sides, U steps ahead,
//     into an LDS ring of RL steps;
//   wave 2 (poller): lane f watches frame line f: the result of each step of
//     that line, polled in the other tile's output (value as flag) UE steps
//     ahead, into the line's halo ring.  A poll that finds the sentinel means
//     the tile has caught up with its producer, so the polls in flight behind
//     it were issued too early as well: they are issued again together.
// Tiles take tickets in dependency order; every wait is bounded (error flag,
// then zeros).
template <int KIND, bool UP, class V>
__global__ __launch_bounds__(192) void stencil_sweep(sweep_args<V> A) {
    using K = sk<KIND>;
    using SH = shape<KIND, UP, V>;
    constexpr int NS = K::NS, TJ = K::TJ, TK = K::TK;
    constexpr int U = SH::U, UE = SH::UE, RL = SH::RL, RG = SH::RG, R = SH::R, RS = SH::RS, NCH = SH::NCH;
    constexpr int NF = frame_count<KIND>();
    constexpr int DM = delta_max<KIND, UP>();
    constexpr unsigned long long USED = frame_used<KIND, UP>();
    static_assert(DM < RG && RL % RG == 0 && R >= RL + DM && NF <= 64, "ring sizes");
    if (A.gate && *A.gate) return;
    __shared__ __align__(16) unsigned char s_rec[RL][NCH * 1024];
    __shared__ V s_bin[RL][64];
    __shared__ V s_halo[(64 + NF) * RS];
    __shared__ int s_tile, s_rdy, s_ext, s_cons;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (threadIdx.x == 0) {
        const unsigned tk = atomicAdd(A.ticket, 1u);
        if (tk < (unsigned)A.ntiles) {
            s_tile = A.tile_of_ticket[tk];
        } else {  // a counter that was not re-armed: flag it, run nothing
            s_tile = -1;
            atomicOr(A.err, 8);
        }
        s_rdy = 0;
        s_ext = 0;
        s_cons = 0;
        if (tk == 0) *A.reset_ticket = 0u;  // the other triangle's counter, for its next sweep
    }
    __syncthreads();
    const int tile = s_tile;
    if (tile < 0) return;
    const int4 m = A.meta[tile];
    const long long base = A.base[tile];
    const int S = m.w;
    const int tfirst = UP ? m.z + S - 1 : m.z;  // global step of local step 0
    const long long last = A.nprow - 1;
    auto clampi = [&](long long e) { return e < 0 ? 0 : (e > last ? last : e); };
    // packed row of this tile's lane at local step s
    auto pidx = [&](int s) { return base + (long long)(UP ? S - 1 - s : s) * 64 + lane; };

    if (wave == 1) {
        // ---------------- streamer
        const u32x4 *rec = reinterpret_cast<const u32x4 *>(A.rec);
        const long long rec0 = base / 64;
        // every load is issued (clamped indices, padded records), so each step adds the same count
        u32x4 ch[U][NCH];
        V bv[U];
        auto issue = [&](int s, int u) {
            const long long g = (rec0 + s) * (SH::REC / 16);
#pragma unroll
            for (int c = 0; c < NCH; c++) ch[u][c] = __builtin_nontemporal_load(rec + g + c * 64 + lane);
            bv[u] = __builtin_nontemporal_load(A.bin + clampi(pidx(s)));
        };
#pragma unroll
        for (int u = 0; u < U; u++) issue(u, u);
        bool bad = false;
        int cons = 0;
        auto consume = [&](int s, int u) {
            // LDS slot s % RL is free once the compute wave has finished step s - RL
            if (!lds_wait(&s_cons, cons, s - RL + 1)) bad = true;
            __asm__ __volatile__("" ::: "memory");  // no staging write moves above the wait
            unsigned char *slot = s_rec[s % RL];
#pragma unroll
            for (int c = 0; c < NCH; c++) *reinterpret_cast<u32x4 *>(slot + c * 1024 + lane * 16) = ch[u][c];
            s_bin[s % RL][lane] = bv[u];
            // LDS operations of one wave complete in order; the wait makes the staged
            // values land before the progress (the loads in flight are not waited for)
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            if (lane == 0) __hip_atomic_store(&s_rdy, s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        int s0 = 0;
        for (; s0 + U <= S; s0 += U) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                consume(s0 + u, u);
                issue(s0 + u + U, u);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (s0 + u < S) consume(s0 + u, u);
        if (__any(bad) && lane == 0) atomicOr(A.err, 1);
        return;
    }

    if (wave == 2) {
        // ---------------- poller: lane f < NF watches frame line f
        const int f = lane < NF ? lane : 0;
        const int jn = m.x + frame_rj<KIND>(f), kn = m.y + frame_rk<KIND>(f);
        const bool watch = lane < NF && ((USED >> f) & 1ull) && jn >= 0 && jn < A.ny && kn >= 0 && kn < A.nz;
        if (!__any(watch)) {  // no frame line in use: nothing to wait for
            if (lane == 0) __hip_atomic_store(&s_ext, S + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return;
        }
        const int tp = watch ? jn / TJ + A.ntj * (kn / TK) : tile;  // the producing tile
        const int4 mp = A.meta[tp];
        const long long bp = A.base[tp];
        const int lanep = watch ? (jn - mp.x) + TJ * (kn - mp.y) : 0;
        const int nbase = tfirst - K::A * jn - K::B * kn;
        auto npos = [&](int s) { return UP ? nbase - s : nbase + s; };  // watched line's row at local step s
        // its packed row: producer base + 64 * (its L step) + its lane
        const long long ib = bp + (long long)(K::A * jn + K::B * kn - mp.z) * 64 + lanep;
        auto eptr = [&](int s) { return A.out + clampi(ib + (long long)npos(s) * 64); };
        auto need = [&](int s) {
            const int p = npos(s);
            return watch && p >= 0 && p < A.nx;
        };
        bool bad = false;
        auto stage = [&](int s, V v) {
            if (lane < NF) s_halo[(64 + lane) * RS + (s & (R - 1))] = need(s) ? v : V(0);
        };
        // results of local steps -DM .. -2: rows before the tile's first step that it reads
#pragma unroll
        for (int d = 2; d <= DM; d++) {
            V v = ld_agent(eptr(-d));
            if (need(-d) && is_sent(v)) v = st_poll(eptr(-d), A.err, bad);
            stage(-d, v);
        }
        // poller step s stages local step s - 1
        V ev[UE];
#pragma unroll
        for (int u = 0; u < UE; u++) ev[u] = ld_agent(eptr(u - 1));
        int cons = 0;
        auto consume = [&](int s, int u) {
            // halo entry s - 1 is free once the compute wave has finished step s - 1 - R + DM
            if (!lds_wait(&s_cons, cons, s - R + DM)) bad = true;
            V v = ev[u];
            const bool miss = need(s - 1) && is_sent(v);
            if (__any(miss)) {
                if (miss) v = st_poll(eptr(s - 1), A.err, bad);
#pragma unroll
                for (int k = 1; k < UE; k++) ev[(u + k) % UE] = ld_agent(eptr(s - 1 + k));
            }
            stage(s - 1, v);
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            if (lane == 0) __hip_atomic_store(&s_ext, s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        int s0 = 0;
        for (; s0 + UE <= S; s0 += UE) {
#pragma unroll
            for (int u = 0; u < UE; u++) {
                consume(s0 + u, u);
                ev[u] = ld_agent(eptr(s0 + u - 1 + UE));
            }
        }
#pragma unroll
        for (int u = 0; u < UE; u++)
            if (s0 + u < S) consume(s0 + u, u);
        if (__any(bad) && lane == 0) atomicOr(A.err, 1);
        return;
    }

    // ---------------- compute
    const int jj = lane % TJ, kk = lane / TJ;
    int hl[NS];  // halo line (times the stride) that slot q reads
#pragma unroll
    for (int q = 0; q < NS; q++) {
        const int rj = jj + sdj<KIND, UP>(q), rk = kk + sdk<KIND, UP>(q);
        const bool in = rj >= 0 && rj < TJ && rk >= 0 && rk < TK;
        hl[q] = (in ? rj + TJ * rk : 64 + frame_id<KIND>(rj, rk)) * RS;
    }
    V *const myhalo = s_halo + lane * RS;
    V ring[RG];  // own results of the last RG steps (the i-1 terms)
#pragma unroll
    for (int u = 0; u < RG; u++) ring[u] = V(0);
    int rdy = 0, ext = 0;
    bool bad = false;
    // bounded wait until a staging wave's counter passes want (cached value first)
    auto wait_for = [&](int *flag, int &seen, int want) {
        for (unsigned spins = 0; seen < want; spins++) {
            seen = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            if (seen >= want) break;
            if (spins > ST_SPIN_LIMIT) {  // never expected: a staging wave stalled
                bad = true;
                seen = want;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __asm__ __volatile__("" ::: "memory");  // no staged read moves above the wait
    };
    // operands of a step, read one step ahead (double buffer b = step & 1), so
    // that a step waits only for its halo values
    V rc[2][NS], rd[2], rb[2];
    unsigned rm[2];
    auto read_rec = [&](int s, int b) {
        const int sl = s % RL;
        const unsigned char *rs = s_rec[sl];
#pragma unroll
        for (int q = 0; q < NS; q++)
            rc[b][q] = *reinterpret_cast<const V *>(rs + q * 64 * (int)sizeof(V) + lane * (int)sizeof(V));
        rm[b] = *reinterpret_cast<const unsigned short *>(rs + SH::OFF_MASK + lane * 2);
        rd[b] = UP ? *reinterpret_cast<const V *>(rs + SH::OFF_DIAG + lane * (int)sizeof(V)) : V(1);
        rb[b] = s_bin[sl][lane];
    };
    wait_for(&s_rdy, rdy, 1);
    read_rec(0, 0);
    for (int s0 = 0; s0 < S; s0 += RG) {
#pragma unroll
        for (int u = 0; u < RG; u++) {
            const int s = s0 + u, b = u & 1;
            if (s < S) {
                // frame values of step s staged; every halo read issued before the first is used
                wait_for(&s_ext, ext, s + 1);
                V nq[NS];
#pragma unroll
                for (int q = 0; q < NS; q++) {
                    const int d = sdelta<KIND, UP>(q);  // 1 <= d <= DM
                    nq[q] = (sdj<KIND, UP>(q) == 0 && sdk<KIND, UP>(q) == 0) ? ring[(u - d + RG) % RG]
                                                                             : s_halo[hl[q] + ((s - d) & (R - 1))];
                }
                // the next step's operands meanwhile
                if (s + 1 < S) {
                    wait_for(&s_rdy, rdy, s + 2);
                    read_rec(s + 1, b ^ 1);
                }
                // the row's terms in CSR order, selected rather than branched around
                V acc = rb[b];
#pragma unroll
                for (int q = 0; q < NS; q++) {
                    const V t = acc - rc[b][q] * nq[q];
                    acc = ((rm[b] >> q) & 1u) ? t : acc;
                }
                if (UP) acc = acc / rd[b];
                ring[u] = acc;
                myhalo[s & (R - 1)] = acc;
                // every lane stores (a lane off the grid or past its line writes a padding row)
                const long long pi = pidx(s);
                st_agent(A.out + pi, acc);
                A.rearm[pi] = sentinel<V>();
                // slot s % RL was read a step ago: released to the streaming wave
                if (lane == 0) __hip_atomic_store(&s_cons, s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
    if (__any(bad) && lane == 0) atomicOr(A.err, 1);
}

// natural order -> packed (padding rows get 0)
template <class TIN, class V>
__global__ void stencil_gather(int64_t nprow, const int *__restrict__ prow, const TIN *__restrict__ x,
                               V *__restrict__ xp, const int *gate) {
    if (gate && *gate) return;
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; p < nprow; p += (int64_t)gridDim.x * blockDim.x) {
        const int r = prow[p];
        xp[p] = r >= 0 ? (V)x[r] : V(0);
    }
}

// packed -> natural order
template <class V, class TOUT>
__global__ void stencil_scatter(int64_t nprow, const int *__restrict__ prow, const V *__restrict__ yp,
                                TOUT *__restrict__ y, const int *gate) {
    if (gate && *gate) return;
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; p < nprow; p += (int64_t)gridDim.x * blockDim.x) {
        const int r = prow[p];
        if (r >= 0) y[r] = (TOUT)yp[p];
    }
}

template <int KIND, class V, class TIN, class TOUT>
void sweep_pair(hala_ilu_s *M, const TIN *x, TOUT *r, const int *gate) {
    hala_ctx c = M->ctx;
    stencil_plan &sp = *M->sp;
    const bool f64 = sizeof(V) == 8;
    V *xp = reinterpret_cast<V *>((f64 ? sp.xp64 : sp.xp32).p);
    V *yp = reinterpret_cast<V *>((f64 ? sp.y64 : sp.y32).p);
    V *up = reinterpret_cast<V *>((f64 ? sp.u64 : sp.u32).p);
    const long long np = sp.L.nprow;
    const int g = grid_for(np, 1024, 8192);
    stencil_gather<TIN, V><<<g, 256, 0, c->stream>>>(np, sp.prow.as<int>(), x, xp, gate);
    HALA_CHECK_LAUNCH();
    {
        stencil_tri &L = sp.L;
        sweep_args<V> a{sp.nx, sp.ny, sp.nz, sp.P, L.ntiles, sp.ntj, np, L.meta.as<int4>(), L.base.as<long long>(),
                        L.tile_of_ticket.as<int>(), L.ticket.as<unsigned>(), sp.U.ticket.as<unsigned>(),
                        (f64 ? L.rec64 : L.rec32).as<unsigned char>(), xp, yp, up, wait_error_flag(c), gate};
        stencil_sweep<KIND, false, V><<<L.ntiles, 192, 0, c->stream>>>(a);
        HALA_CHECK_LAUNCH();
    }
    {
        stencil_tri &Up = sp.U;
        sweep_args<V> a{sp.nx, sp.ny, sp.nz, sp.P, Up.ntiles, sp.ntj, np, Up.meta.as<int4>(), Up.base.as<long long>(),
                        Up.tile_of_ticket.as<int>(), Up.ticket.as<unsigned>(), sp.L.ticket.as<unsigned>(),
                        (f64 ? Up.rec64 : Up.rec32).as<unsigned char>(), yp, up, yp, wait_error_flag(c), gate};
        stencil_sweep<KIND, true, V><<<Up.ntiles, 192, 0, c->stream>>>(a);
        HALA_CHECK_LAUNCH();
    }
    stencil_scatter<V, TOUT><<<g, 256, 0, c->stream>>>(np, sp.prow.as<int>(), up, r, gate);
    HALA_CHECK_LAUNCH();
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
    if (!up) {
        sp.prow.reserve((size_t)acc * sizeof(int));
        pa.prow = sp.prow.as<int>();
    }
    dim3 g((unsigned)((maxS * 64 + 255) / 256), (unsigned)ntiles);
    stencil_pack<<<g, 256, 0, c->stream>>>(pa);
    HALA_CHECK_LAUNCH();
}

}  // namespace

// The packed vectors of one precision: input, L's output and U's output, both
// outputs all-sentinel (each sweep re-arms the other's rows as it goes).
static void packed_vectors(hala_ilu_s *M, bool f32) {
    hala_ctx c = M->ctx;
    stencil_plan &sp = *M->sp;
    const int64_t np = sp.L.nprow;
    const size_t vs = f32 ? sizeof(float) : sizeof(double);
    dbuf &xp = f32 ? sp.xp32 : sp.xp64, &y = f32 ? sp.y32 : sp.y64, &u = f32 ? sp.u32 : sp.u64;
    xp.reserve((size_t)np * vs);
    y.reserve((size_t)np * vs);
    u.reserve((size_t)np * vs);
    const int g = grid_for(np, 1024, 8192);
    if (f32) {
        fill_sent<float><<<g, 256, 0, c->stream>>>(y.as<float>(), np);
        fill_sent<float><<<g, 256, 0, c->stream>>>(u.as<float>(), np);
    } else {
        fill_sent<double><<<g, 256, 0, c->stream>>>(y.as<double>(), np);
        fill_sent<double><<<g, 256, 0, c->stream>>>(u.as<double>(), np);
    }
    HALA_CHECK_LAUNCH();
}

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
        packed_vectors(M, false);
        return true;
    }
    return false;
}

// bytes of one step record (the layout of shape<>::REC)
static int record_bytes(int NS, bool up, int vsize) { return NS * 64 * vsize + (up ? 64 * vsize : 0) + 128; }

void stencil_load(hala_ilu_s *M, bool f32) {
    hala_ctx c = M->ctx;
    stencil_plan &sp = *M->sp;
    const kind_info K = info_of(sp.kind);
    const double *f = M->factor.as<double>();
    auto load = [&](stencil_tri &T, bool up) {
        const int64_t nrec = T.nprow / 64;
        auto fill = [&](dbuf &rec, auto zero) {
            using V = decltype(zero);
            const int rb = record_bytes(K.NS, up, (int)sizeof(V));
            // the streaming wave reads up to 16 records past a tile and 1 KiB past a record
            rec.reserve((size_t)(nrec + 16) * rb + 1024);
            stencil_records<V><<<grid_for(nrec * 64, 1024, 8192), 256, 0, c->stream>>>(
                nrec, K.NS, up ? 1 : 0, rb, T.pos.as<int>(), up ? T.dpos.as<int>() : nullptr,
                T.mask.as<unsigned short>(), f, rec.as<unsigned char>());
            HALA_CHECK_LAUNCH();
        };
        fill(T.rec64, 0.0);
        if (f32) fill(T.rec32, 0.0f);
    };
    load(sp.L, false);
    load(sp.U, true);
    if (f32 && !sp.y32.p) packed_vectors(M, true);
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
