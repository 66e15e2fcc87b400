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
// that the loader wave moves a step's operands with NCH 16-byte loads per lane.
template <int KIND, bool UP, class V>
struct shape {
    static constexpr int NS = sk<KIND>::NS;
    static constexpr int OFF_DIAG = NS * 64 * (int)sizeof(V);
    static constexpr int OFF_MASK = OFF_DIAG + (UP ? 64 * (int)sizeof(V) : 0);
    static constexpr int REC = OFF_MASK + 128;       // bytes per step record (a multiple of 128)
    static constexpr int NCH = (REC + 1023) / 1024;  // 16-byte loads per lane and step
    // steps of loads in flight: at most 56 of the wave's 63 outstanding loads
    static constexpr int U = (NCH + 2) * 8 <= 56 ? 8 : 4;
    static constexpr int RL = 8;   // staged steps in LDS = compute unroll = own-result ring
    static constexpr int R = 16;   // halo ring (steps), a power of two
    static constexpr int RS = R + 1;  // halo line stride (odd: lanes of different lines hit different banks)
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

template <class V, class TB, class TO, class TR>
struct sweep_args {
    int nx, ny, nz, P, ntiles, n;
    const int4 *meta;
    const long long *base;
    const int *tile_of_ticket;
    unsigned *ticket, *reset_ticket;
    const unsigned char *rec;  // the triangle's step records in precision V
    const TB *bin;             // right-hand side, natural order
    TO *out;                   // result, natural order (polled by the other tiles' loaders)
    TR *rearm;                 // gets the sentinel at each row this sweep computes
    int *err;
    const int *gate;
};

constexpr unsigned ST_SPIN_LIMIT = 1u << 22;

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Polls a result until its producer has written it.  Loads only: on the spin
// limit, or when another wave has raised the error flag, sets bad and returns 0.
// Every load it issues is examined before it returns, so no load of the poll
// is left in flight for the caller's waits to count.
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

// Two waves per tile.  The loader wave issues every global load of the tile
// and nothing else (no store, no atomic until its last step), so the
// compiler's wait for a step's loads counts only the later steps' loads and
// the U steps in flight stay in flight; the compute wave reads LDS only and
// stores the results.
//
// LDS: s_rec / s_bin: RL staged steps (records and right-hand sides); s_halo:
// for each of the tile's 64 lines and of the frame lines, a ring of the last R
// results by local step.  The compute wave writes its lines' entries, the
// loader the frame lines' entries (polled in global memory: value as flag).
template <int KIND, bool UP, class V, class TB, class TO, class TR>
__global__ __launch_bounds__(128) void stencil_sweep(sweep_args<V, TB, TO, TR> A) {
    using K = sk<KIND>;
    using SH = shape<KIND, UP, V>;
    constexpr int NS = K::NS, TJ = K::TJ, TK = K::TK;
    constexpr int U = SH::U, RL = SH::RL, R = SH::R, RS = SH::RS, NCH = SH::NCH;
    constexpr int NF = frame_count<KIND>();
    constexpr int DM = delta_max<KIND, UP>();
    constexpr unsigned long long USED = frame_used<KIND, UP>();
    static_assert(DM < RL && R >= RL + DM && NF <= 64, "ring sizes");
    if (A.gate && *A.gate) return;
    __shared__ __align__(16) unsigned char s_rec[RL][NCH * 1024];
    __shared__ V s_bin[RL][64];
    __shared__ V s_halo[(64 + NF) * RS];
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
    const long long rec0 = A.base[tile] / 64;  // the tile's first step record
    const int S = m.w;
    const int jj = lane % TJ, kk = lane / TJ;
    const int j = m.x + jj, k = m.y + kk;
    const int tfirst = UP ? m.z + m.w - 1 : m.z;
    const int ibase = tfirst - K::A * j - K::B * k;  // i = ibase + s (L), ibase - s (U)
    const long long rowbase = (long long)A.nx * j + (long long)A.P * k;

    if (wave == 1) {
        // ---------------- loader: lane f < NF also watches frame line f
        const int f = lane < NF ? lane : 0;
        const int jn = m.x + frame_rj<KIND>(f), kn = m.y + frame_rk<KIND>(f);
        const bool watch = lane < NF && ((USED >> f) & 1ull) && jn >= 0 && jn < A.ny && kn >= 0 && kn < A.nz;
        const int nbase = tfirst - K::A * jn - K::B * kn;
        const long long nrow = (long long)A.nx * jn + (long long)A.P * kn;
        const long long last = (long long)A.n - 1;
        auto clampi = [&](long long e) { return e < 0 ? 0 : (e > last ? last : e); };
        auto npos = [&](int s) { return UP ? nbase - s : nbase + s; };  // watched line's row at local step s
        const u32x4 *rec = reinterpret_cast<const u32x4 *>(A.rec);
        // loads of U steps in flight; every load is issued (clamped addresses, padded
        // records), so each step adds the same count
        u32x4 ch[U][NCH];
        TO ex[U];
        TB bv[U];
        auto issue = [&](int s, int u) {
            const long long g = (rec0 + s) * (SH::REC / 16);
#pragma unroll
            for (int c = 0; c < NCH; c++) ch[u][c] = __builtin_nontemporal_load(rec + g + c * 64 + lane);
            ex[u] = ld_agent(A.out + clampi(nrow + npos(s - 1)));
            bv[u] = A.bin[clampi(rowbase + (UP ? ibase - s : ibase + s))];
        };
        bool bad = false;
        // the watched line's result of local step s into its halo ring
        auto stage_line = [&](int s, TO v) {
            const int p = npos(s);
            const bool need = watch && p >= 0 && p < A.nx;
            if (need && is_sent(v)) v = st_poll(A.out + (nrow + p), A.err, bad);
            if (lane < NF) s_halo[(64 + lane) * RS + (s & (R - 1))] = need ? (V)v : V(0);
        };
        // frame results of local steps -DM .. -2: rows before the tile's first step that it reads
#pragma unroll
        for (int d = 2; d <= DM; d++) stage_line(-d, ld_agent(A.out + clampi(nrow + npos(-d))));
#pragma unroll
        for (int u = 0; u < U; u++) issue(u, u);
        int cons = 0;
        auto consume = [&](int s, int u) {
            // LDS slot s % RL is free once the compute wave has finished step s - RL
            for (unsigned spins = 0; cons < s - RL + 1; spins++) {
                cons = __hip_atomic_load(&s_cons, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (cons >= s - RL + 1) break;
                if (spins > ST_SPIN_LIMIT) {  // never expected: the compute wave stalled
                    bad = true;
                    cons = s;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __asm__ __volatile__("" ::: "memory");  // no staging write moves above the check
            unsigned char *slot = s_rec[s % RL];
#pragma unroll
            for (int c = 0; c < NCH; c++) *reinterpret_cast<u32x4 *>(slot + c * 1024 + lane * 16) = ch[u][c];
            s_bin[s % RL][lane] = (V)bv[u];
            stage_line(s - 1, ex[u]);
            // LDS operations of one wave complete in order; the wait makes the staged
            // values land before the progress (the loads in flight are not waited for)
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            if (lane == 0) __hip_atomic_store(&s_ready, s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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

    // ---------------- compute
    const bool active = j < A.ny && k < A.nz;
    int hl[NS];  // halo line (times the stride) that slot q reads
#pragma unroll
    for (int q = 0; q < NS; q++) {
        const int rj = jj + sdj<KIND, UP>(q), rk = kk + sdk<KIND, UP>(q);
        const bool in = rj >= 0 && rj < TJ && rk >= 0 && rk < TK;
        hl[q] = (in ? rj + TJ * rk : 64 + frame_id<KIND>(rj, rk)) * RS;
    }
    V *const myhalo = s_halo + lane * RS;
    V ring[RL];  // own results of the last RL steps (the i-1 terms)
#pragma unroll
    for (int u = 0; u < RL; u++) ring[u] = V(0);
    int rdy = 0;
    for (int s0 = 0; s0 < S; s0 += RL) {
#pragma unroll
        for (int u = 0; u < RL; u++) {
            const int s = s0 + u;
            if (s < S) {
                for (unsigned spins = 0; rdy <= s; spins++) {
                    rdy = __hip_atomic_load(&s_ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (rdy > s) break;
                    if (spins > ST_SPIN_LIMIT) {  // never expected: the loader stalled
                        atomicOr(A.err, 1);
                        rdy = s + 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __asm__ __volatile__("" ::: "memory");  // no staged read moves above the check
                const unsigned char *rs = s_rec[u];
                V acc = s_bin[u][lane];
                const unsigned msk = *reinterpret_cast<const unsigned short *>(rs + SH::OFF_MASK + lane * 2);
#pragma unroll
                for (int q = 0; q < NS; q++) {
                    const V c = *reinterpret_cast<const V *>(rs + q * 64 * (int)sizeof(V) + lane * (int)sizeof(V));
                    const int d = sdelta<KIND, UP>(q);  // 1 <= d <= DM
                    V nv;
                    if (sdj<KIND, UP>(q) == 0 && sdk<KIND, UP>(q) == 0) nv = ring[(u - d + RL) % RL];
                    else nv = s_halo[hl[q] + ((s - d) & (R - 1))];
                    if ((msk >> q) & 1u) acc = acc - c * nv;
                }
                if (UP) acc = acc / *reinterpret_cast<const V *>(rs + SH::OFF_DIAG + lane * (int)sizeof(V));
                ring[u] = acc;
                myhalo[s & (R - 1)] = acc;
                const int i = UP ? ibase - s : ibase + s;
                if (active && i >= 0 && i < A.nx) {
                    st_agent(A.out + (rowbase + i), (TO)acc);
                    A.rearm[rowbase + i] = sentinel<TR>();
                }
                __asm__ __volatile__("" ::: "memory");  // the staged reads stay before the release of their slot
                if (lane == 0) __hip_atomic_store(&s_cons, s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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
        sweep_args<V, TIN, V, TOUT> a{sp.nx, sp.ny, sp.nz, sp.P, L.ntiles, M->n, L.meta.as<int4>(),
                                      L.base.as<long long>(), L.tile_of_ticket.as<int>(), L.ticket.as<unsigned>(),
                                      sp.U.ticket.as<unsigned>(), (f64 ? L.rec64 : L.rec32).as<unsigned char>(),
                                      x, y, r, wait_error_flag(c), gate};
        stencil_sweep<KIND, false, V, TIN, V, TOUT><<<L.ntiles, 128, 0, c->stream>>>(a);
        HALA_CHECK_LAUNCH();
    }
    {
        stencil_tri &Up = sp.U;
        sweep_args<V, V, TOUT, V> a{sp.nx, sp.ny, sp.nz, sp.P, Up.ntiles, M->n, Up.meta.as<int4>(),
                                    Up.base.as<long long>(), Up.tile_of_ticket.as<int>(), Up.ticket.as<unsigned>(),
                                    sp.L.ticket.as<unsigned>(), (f64 ? Up.rec64 : Up.rec32).as<unsigned char>(),
                                    y, r, y, wait_error_flag(c), gate};
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
            // the loader reads up to 8 records past a tile and 1 KiB past a record
            rec.reserve((size_t)(nrec + 8) * rb + 1024);
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
