// common.hpp -- shared host/device plumbing of libhala_amd.so (gfx950 only).
//
// Arithmetic contract: the library is compiled with -ffp-contract=off, so a
// product followed by a sum is never fused; every reduction follows the
// canonical order documented at tree256() below.  Together with sequential
// per-row sums this makes every FP64 result reproducible run to run.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

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
struct dbuf {
    void *p = nullptr;
    size_t bytes = 0;
    dbuf() = default;
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    ~dbuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // grows (never shrinks); contents are not preserved
    void reserve(size_t b) {
        if (b <= bytes) return;
        release();
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            throw status_error{HALA_ERR_ALLOC, "hipMalloc of " + std::to_string(b) + " bytes failed"};
        }
        bytes = b;
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
struct hala_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cus = 256;
    // small device scratch: reduction partials, scalars, counters
    hala_amd::dbuf red;      // reduction partials
    hala_amd::dbuf scal;     // device scalars (8 KiB), zero-initialised at creation
    hala_amd::dbuf vec;      // solver work vectors
    double *host_pinned = nullptr;  // pinned host landing zone for scalars
};

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

// Final canonical pass over nb partials by one 256-thread block:
// s[t] = 0 + p[t] + p[t+256] + ..., then tree256.  Result valid in thread 0.
template <class T>
__device__ __forceinline__ T final_pass(const T *part, int64_t nb, T *sh) {
    T s = T(0);
    for (int64_t k = threadIdx.x; k < nb; k += CB) s = s + ld_agent(part + k);
    return tree256_1<T>(s, sh);
}

// "Last block" completion: every block calls this after writing its partials
// (thread 0 only stores them).  Returns true in exactly one block, after all
// partials are visible; the counter is re-armed to zero for the next launch.
__device__ __forceinline__ bool last_block(unsigned *counter, unsigned nblocks) {
    __shared__ int am_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        am_last = (prev == nblocks - 1);
        if (am_last) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return am_last != 0;
}

inline unsigned *counters(hala_ctx c) { return reinterpret_cast<unsigned *>(c->scal.as<double>() + S_CNT); }

inline int grid_for(int64_t work, int per_block, int cap) {
    int64_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace hala_amd
