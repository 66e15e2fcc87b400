// common.hpp -- shared host/device plumbing of libhala_amd.so (gfx950 only).
//
// Arithmetic contract: the library is compiled with -ffp-contract=off, so a
// product followed by a sum is never fused; every reduction follows the
// canonical order documented at tree256() below.  Together with sequential
// per-row sums this makes every FP64 result reproducible run to run.
#pragma once
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hala_amd.h"

namespace hala_amd {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);

struct status_error {
    int code;
    std::string msg;
};

#define HALA_HIP(call)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::hala_amd::status_error{HALA_ERR_HIP, std::string(#call) + ": " +           \
                                                             hipGetErrorString(e_)};           \
    } while (0)

#define HALA_REQUIRE(cond, code, msg)                                                          \
    do {                                                                                       \
        if (!(cond)) throw ::hala_amd::status_error{code, std::string(msg)};                   \
    } while (0)

#define HALA_CHECK_LAUNCH() HALA_HIP(hipGetLastError())

// A roctx range over one entry point of the C ABI (named by the entry point),
// so a trace groups the kernels of each call: rocprofv3 --marker-trace.
struct trace_range {
    explicit trace_range(const char *name) { roctxRangePushA(name); }
    ~trace_range() { roctxRangePop(); }
    trace_range(const trace_range &) = delete;
    trace_range &operator=(const trace_range &) = delete;
};
#define HALA_RANGE() ::hala_amd::trace_range hala_range_(__func__)

// Runs body, converting exceptions to status codes and the last-error string.
template <class F>
int guard(F &&body) {
    try {
        body();
        return HALA_OK;
    } catch (status_error &e) {
        set_error(e.msg);
        return e.code;
    } catch (std::bad_alloc &) {
        set_error("host allocation failed");
        return HALA_ERR_ALLOC;
    } catch (std::exception &e) {
        set_error(e.what());
        return HALA_ERR_INTERNAL;
    }
}

// ---------------------------------------------------------------------------
// device buffers owned by the library
// ---------------------------------------------------------------------------
// Device blocks kept for reuse by one context: an ILU analysis allocates some
// hundred buffers and its destruction frees them, and each hipMalloc / hipFree
// costs tens to hundreds of microseconds (hipFree also waits for the whole
// device).  Inside an alloc_scope, buffers are taken from and returned to the
// scope's cache.  A cache serves one stream: a block released while kernels
// on that stream may still read it is handed out again only to later work on
// the same stream, and owners synchronise the stream before their final
// release (an ILU destroy does).
struct buf_cache {
    std::mutex m;
    std::multimap<size_t, void *> blocks;  // size -> block
    size_t held = 0;
    static constexpr size_t CAP = size_t(16) << 30;  // bytes kept at most
    // a block of at least b bytes and at most 2b, or nullptr
    void *take(size_t b, size_t *got) {
        std::lock_guard<std::mutex> g(m);
        auto it = blocks.lower_bound(b);
        if (it == blocks.end() || it->first > 2 * b) return nullptr;
        void *p = it->second;
        *got = it->first;
        held -= it->first;
        blocks.erase(it);
        return p;
    }
    void put(void *p, size_t b) {
        {
            std::lock_guard<std::mutex> g(m);
            if (held + b <= CAP) {
                blocks.emplace(b, p);
                held += b;
                return;
            }
        }
        (void)hipFree(p);
    }
    ~buf_cache() {
        for (auto &kv : blocks) (void)hipFree(kv.second);
    }
};

inline thread_local buf_cache *tl_cache = nullptr;

struct alloc_scope {
    buf_cache *c0;
    explicit alloc_scope(buf_cache *c) : c0(tl_cache) { tl_cache = c; }
    ~alloc_scope() { tl_cache = c0; }
    alloc_scope(const alloc_scope &) = delete;
    alloc_scope &operator=(const alloc_scope &) = delete;
};

struct dbuf {
    void *p = nullptr;
    size_t bytes = 0;
    buf_cache *cache = nullptr;  // where the block returns (nullptr: hipFree)
    dbuf() = default;
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    dbuf(dbuf &&o) noexcept : p(o.p), bytes(o.bytes), cache(o.cache) {
        o.p = nullptr;
        o.bytes = 0;
        o.cache = nullptr;
    }
    dbuf &operator=(dbuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            cache = o.cache;
            o.p = nullptr;
            o.bytes = 0;
            o.cache = nullptr;
        }
        return *this;
    }
    void swap(dbuf &o) noexcept {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(cache, o.cache);
    }
    ~dbuf() { release(); }
    void release() {
        if (p) {
            if (cache) cache->put(p, bytes);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        cache = nullptr;
    }
    // grows (never shrinks); contents are not preserved
    void reserve(size_t b) {
        if (b <= bytes) return;
        release();
        size_t got = b;
        if (tl_cache) p = tl_cache->take(b, &got);
        if (!p) {
            hipError_t e = hipMalloc(&p, b);
            if (e != hipSuccess) {
                p = nullptr;
                throw status_error{HALA_ERR_ALLOC, "hipMalloc of " + std::to_string(b) + " bytes failed"};
            }
            got = b;
        }
        cache = tl_cache;
        bytes = got;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

}  // namespace hala_amd

// ---------------------------------------------------------------------------
// context (the engine): device, stream and reusable scratch
// ---------------------------------------------------------------------------
// Launch-shape options (set with hala_ctx_set_option, picked by the autotuner).
struct hala_options {
    int tri_waves_per_cu = 8;  // persistent triangular-solve pool: waves per CU
    int tri_pool_waves = 0;    // absolute pool size (overrides waves per CU when > 0)
    int tri_sleep = 1;         // s_sleep slots between polls of an unready dependency
    int cg_batch_max = 32;     // solver passes launched between stop-flag reads
    int spmv_blocks_per_cu = 0;  // fused SpMV + dot: workgroups per CU of a grid-stride launch (0: one per block)
    int tri_streams = 1;       // ticket counters of the triangular-solve pool (1..TK_MAX)
    int tri_col = 1;           // column-blocked triangular sweeps where the pattern allows (0: level pool only)
    int tri_col_cols = 0;      // columns of the blocked sweep (0: smallest that fits)
    int tri_col_split = 0;     // columns per slab of planes (0: automatic, 1: contiguous row blocks)
    int cg_fuse = 1;           // ILU-PCG: the sweeps' gather/scatter fused into the update and r.z kernels
    int cg_xcd = 1;            // fused ILU-PCG: the update (scatter into L's slot order) in XCD-aware block order
    int ilu32_jacobi = 0;      // FP32 ILU applies between FP64 vectors: k > 0 Jacobi sweeps per triangle (0: exact)
    int ilu64_jacobi = 0;      // FP64 ILU applies inside GMRES and hala_ilu_apply_f64: k > 0 Jacobi sweeps per
                               // triangle (0: exact); CG keeps the exact sweeps
    int tri_stencil = 0;       // 1: ILU(0) sweeps of a stencil pattern on the pencil schedule (0: general sweeps)
    int sync_checks = 1;       // standalone sweeps (ILU apply, trsv) wait for the device and report a wait that
                               // gave up before returning; 0: they return at once and such an error is reported
                               // by the next call that synchronises the context (hala_ctx_sync included)
};

struct hala_ctx_s {
    hala_options opt;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cus = 256;
    // small device scratch: reduction partials, scalars, counters
    hala_amd::dbuf red;      // reduction partials, FP64 (sentinel between reductions)
    hala_amd::dbuf red32;    // reduction partials, FP32 (sentinel between reductions)
    hala_amd::dbuf scal;     // device scalars (8 KiB), zero-initialised at creation
    hala_amd::dbuf vec;      // solver work vectors
    hala_amd::dbuf lbc;      // ticket counters of the triangular-solve pool (zeroed at creation, self re-arming)
    double *host_pinned = nullptr;  // pinned host landing zone for scalars
    double *noalloc = nullptr;      // this device's scratch of the allocation-free entry point
    // temporaries of the analysis passes (see tmp_scope)
    std::vector<std::unique_ptr<hala_amd::dbuf>> tmp;
    size_t tmp_top = 0;
    // second context on a stream of its own, for analysis work run beside this
    // context's (created on first use, destroyed with this context)
    hala_ctx_s *aux = nullptr;
    // blocks of the ILU objects built on this context, kept for the next ones
    std::unique_ptr<hala_amd::buf_cache> cache{new hala_amd::buf_cache()};
};

namespace hala_amd {
// Temporary device buffers of one library call, kept by the context and reused
// by later calls instead of allocated and freed each time (hipFree waits for
// the whole device).  Every user runs on the context's stream, so a reuse is
// ordered after the previous one.  Scopes nest (last in, first out).
class tmp_scope {
  public:
    explicit tmp_scope(hala_ctx_s *c) : c_(c), mark_(c->tmp_top) {}
    ~tmp_scope() { c_->tmp_top = mark_; }
    tmp_scope(const tmp_scope &) = delete;
    tmp_scope &operator=(const tmp_scope &) = delete;
    dbuf &get() {
        if (c_->tmp_top == c_->tmp.size()) c_->tmp.emplace_back(new dbuf());
        return *c_->tmp[c_->tmp_top++];
    }

  private:
    hala_ctx_s *c_;
    size_t mark_;
};
}  // namespace hala_amd

namespace hala_amd {

constexpr int CB = 256;  // canonical block (elements per partial, width of the final pass)

// Layout of hala_ctx_s::scal (doubles / unsigned counters)
enum : int {
    S_DOT = 0,       // result of a standalone reduction
    S_CNT = 64,      // unsigned counters start here (as unsigned*, index in units of unsigned)
    S_SOLVER = 128,  // solver scalars
    S_TOTAL = 1024
};

inline void set_device(hala_ctx c) { HALA_HIP(hipSetDevice(c->device)); }

// The context's auxiliary context (same device, options copied, its own
// non-blocking stream of default priority), created on first use.
hala_ctx aux_ctx(hala_ctx c);

// Bit patterns never produced by arithmetic: "not yet computed" markers for
// value-as-flag dependency waits (a signalling NaN with a fixed payload).
constexpr uint64_t SENT64 = 0x7FF4DEADBEEF0001ull;
constexpr uint32_t SENT32 = 0x7FA0BEEFu;

__device__ __forceinline__ double sentinel64() { return __longlong_as_double((long long)SENT64); }
__device__ __forceinline__ float sentinel32() { return __uint_as_float(SENT32); }
__device__ __forceinline__ bool is_sent(double v) { return (uint64_t)__double_as_longlong(v) == SENT64; }
__device__ __forceinline__ bool is_sent(float v) { return __float_as_uint(v) == SENT32; }

template <class T>
__device__ __forceinline__ T sentinel();
template <>
__device__ __forceinline__ double sentinel<double>() {
    return sentinel64();
}
template <>
__device__ __forceinline__ float sentinel<float>() {
    return sentinel32();
}

// Agent-scope relaxed accesses: coherent across the XCDs without fences.
template <class T>
__device__ __forceinline__ T ld_agent(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_agent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Canonical tree over one 256-thread block: a[t] += a[t+off] for
// off = 128, 64, ..., 1; the total lands in thread 0.  Offsets 128 and 64
// cross wavefronts through LDS; 32..1 stay in wavefront 0 as lane shuffles,
// which add exactly the same operand pairs.  Requires blockDim.x == 256 and
// sh sized 256*K.
template <class T, int K>
__device__ __forceinline__ void tree256(T (&v)[K], T *sh) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; k++) sh[k * CB + t] = v[k];
    __syncthreads();
    if (t < 128) {
#pragma unroll
        for (int k = 0; k < K; k++) sh[k * CB + t] = sh[k * CB + t] + sh[k * CB + t + 128];
    }
    __syncthreads();
    if (t < 64) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            T a = sh[k * CB + t] + sh[k * CB + t + 64];
            a = a + __shfl_down(a, 32, 64);
            a = a + __shfl_down(a, 16, 64);
            a = a + __shfl_down(a, 8, 64);
            a = a + __shfl_down(a, 4, 64);
            a = a + __shfl_down(a, 2, 64);
            a = a + __shfl_down(a, 1, 64);
            v[k] = a;
        }
    }
    __syncthreads();
}

template <class T>
__device__ __forceinline__ T tree256_1(T v, T *sh) {
    T a[1] = {v};
    tree256<T, 1>(a, sh);
    return a[0];
}

// Partials of a reduction are value-as-flag slots: the buffer holds sentinels
// between reductions; each block stores its partials with agent-scope stores,
// and the collecting block polls every slot until it is no longer the
// sentinel, then re-arms it.  No release/acquire fence is needed (on gfx950 an
// agent-scope fence writes back and invalidates the whole L2, which costs far
// more than the reduction itself).
//
// A wait is bounded: after 2^24 polls (seconds; never expected) it raises the
// context's wait-error flag `err` and the partial counts as 0.  The host reads
// the flag before it reports a result (check_wait_error), throws, and re-arms
// every partial slot once the device is idle, so a partial that lands late
// cannot be mistaken for a later reduction's.
constexpr unsigned RED_SPIN_LIMIT = 1u << 24;

// Final canonical pass over nb partials by one 256-thread block:
// s[t] = 0 + p[t] + p[t+256] + ..., then tree256.  Result valid in thread 0.
// Re-arms every slot it reads.  The slots of a batch still holding the
// sentinel are re-polled together, so partials that land late cost one round
// trip after the last of them rather than one each.  A wait is bounded: after
// RED_SPIN_LIMIT polls it raises bit 4 of err and the partials still missing
// count as 0.
template <class T>
__device__ __forceinline__ T final_pass(T *part, int64_t nb, T *sh, int *err) {
    constexpr int B = 8;  // slots loaded per thread at once
    T s = T(0);
    for (int64_t k0 = threadIdx.x; k0 < nb; k0 += (int64_t)CB * B) {
        T v[B];
        unsigned pend = 0;
#pragma unroll
        for (int j = 0; j < B; j++) {
            const int64_t k = k0 + (int64_t)CB * j;
            v[j] = (k < nb) ? ld_agent(part + k) : T(0);
            if (k < nb && is_sent(v[j])) pend |= 1u << j;
        }
        for (unsigned spins = 0; pend != 0; spins++) {
            if (spins > RED_SPIN_LIMIT) {
                atomicOr(err, 4);
#pragma unroll
                for (int j = 0; j < B; j++)
                    if (pend >> j & 1u) v[j] = T(0);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int j = 0; j < B; j++)
                if (pend >> j & 1u) {
                    const T x = ld_agent(part + k0 + (int64_t)CB * j);
                    if (!is_sent(x)) {
                        v[j] = x;
                        pend &= ~(1u << j);
                    }
                }
        }
#pragma unroll
        for (int j = 0; j < B; j++) {
            const int64_t k = k0 + (int64_t)CB * j;
            if (k < nb) {
                st_agent(part + k, sentinel<T>());
                s = s + v[j];
            }
        }
    }
    return tree256_1<T>(s, sh);
}

// Collector election: every block calls this after its thread 0 stored its
// partials; it is true in block 0 (of each grid row), which then collects the
// partials with final_pass().  The partials are value-as-flag slots, so the
// collector needs no arrival count: final_pass() waits on each slot until its
// block has stored it.  Block 0 is dispatched first and the other blocks never
// wait on anything, so its wait always ends; the critical path after the last
// store is one poll round trip, with no chain of atomics in front of it.
__device__ __forceinline__ bool is_collector() { return blockIdx.x == 0; }

// Reduction partial buffers (sentinel-filled when grown), defined in ctx.hip.
double *partials64(hala_ctx c, size_t count);
float *partials32(hala_ctx c, size_t count);

inline unsigned *counters(hala_ctx c) { return reinterpret_cast<unsigned *>(c->scal.as<double>() + S_CNT); }

// Device error flag raised by a dependency or reduction wait that gave up
// (bit 1: a triangular-solve wait, 2: a pool that no wave ran, 4: a reduction
// partial).  Read and cleared by check_wait_error (trsolve.hip).
inline int *wait_error_flag(hala_ctx c) { return reinterpret_cast<int *>(counters(c) + 120); }

// Throws HALA_ERR_NOCONV if a wait on the context's stream gave up since the
// last check, after clearing the flag and, once the stream is idle, re-arming
// every reduction partial slot (a partial stored after its collector gave up
// would otherwise be read by the next reduction).  Synchronises the stream.
void check_wait_error(hala_ctx c);
// The flag word as the host last copied it with the stream synchronised
// (host_pinned must hold the scalar block prefix up to the flag).
void check_wait_error_host(hala_ctx c, int flag);

// Ticket counters of the triangular-solve pool: TK_MAX counters plus one
// "waves finished" counter, each on its own 4 KiB page so that their atomics
// land in different memory channels instead of queueing on one address.
constexpr int TK_STRIDE = 1024;  // in unsigned words
constexpr int TK_MAX = 4;
constexpr int LBC_WORDS = (TK_MAX + 1) * TK_STRIDE;
inline unsigned *tri_tickets(hala_ctx c) { return c->lbc.as<unsigned>(); }
constexpr int GM_MAX_RESTART = 64;

// XCD-aware block order.  Workgroups are dispatched round-robin over the
// chip's 8 XCDs (workgroup b runs on XCD b % 8), each XCD with its own L2.
// A kernel launched with xcd_grid(nvirt) workgroups handles virtual block
// xcd_block(nvirt) (>= nvirt: nothing): every XCD takes a contiguous eighth
// of the virtual blocks, so blocks that touch the same lines share an L2.
constexpr int NXCD = 8;
inline int xcd_grid(int nvirt) { return (nvirt + NXCD - 1) / NXCD * NXCD; }
__device__ __forceinline__ int xcd_block(int nvirt) {
    const int per = (nvirt + NXCD - 1) / NXCD;
    return (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
}

inline int grid_for(int64_t work, int per_block, int cap) {
    int64_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace hala_amd
