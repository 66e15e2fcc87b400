// ctx.hip -- engine (device + stream), device memory, fills and conversions.
#include "common.hpp"

namespace hala_amd {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

template <class T>
__global__ void fill_kernel(T *__restrict__ x, int64_t n, T v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) x[i] = v;
}

template <class A, class B>
__global__ void convert_kernel(const A *__restrict__ x, B *__restrict__ y, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) y[i] = (B)x[i];
}

template <class T>
static void fill(hala_ctx c, T *x, int64_t n, T v) {
    HALA_REQUIRE(c && n >= 0 && (x || n == 0), HALA_ERR_ARG, "fill: bad argument");
    if (n == 0) return;
    set_device(c);
    int g = grid_for(n, 256 * 4, 8192);
    fill_kernel<T><<<g, 256, 0, c->stream>>>(x, n, v);
    HALA_CHECK_LAUNCH();
}

template <class A, class B>
static void convert(hala_ctx c, int64_t n, const A *x, B *y) {
    HALA_REQUIRE(c && n >= 0 && ((x && y) || n == 0), HALA_ERR_ARG, "convert: bad argument");
    if (n == 0) return;
    set_device(c);
    int g = grid_for(n, 256 * 4, 8192);
    convert_kernel<A, B><<<g, 256, 0, c->stream>>>(x, y, n);
    HALA_CHECK_LAUNCH();
}

template <class T>
__global__ void fill_sentinel_kernel(T *__restrict__ x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) x[i] = sentinel<T>();
}

template <class T>
static T *partials(hala_ctx c, dbuf &b, size_t count) {
    if (count < 1) count = 1;
    size_t need = count * sizeof(T);
    if (need > b.bytes) {
        HALA_HIP(hipStreamSynchronize(c->stream));  // a reduction in flight may still read the old buffer
        b.reserve(need < (size_t)(64 * 1024) * sizeof(T) ? (size_t)(64 * 1024) * sizeof(T) : need);
        int64_t n = (int64_t)(b.bytes / sizeof(T));
        fill_sentinel_kernel<T><<<grid_for(n, 256 * 4, 4096), 256, 0, c->stream>>>(b.as<T>(), n);
        HALA_CHECK_LAUNCH();
    }
    return b.as<T>();
}

double *partials64(hala_ctx c, size_t count) { return partials<double>(c, c->red, count); }
float *partials32(hala_ctx c, size_t count) { return partials<float>(c, c->red32, count); }

int fetch_i32(hala_ctx c, const int *p);
double *noalloc_scratch(int device);

void check_wait_error_host(hala_ctx c, int flag) {
    if (flag == 0) return;
    HALA_HIP(hipMemsetAsync(wait_error_flag(c), 0, sizeof(int), c->stream));
    if (flag & 4) {
        // a collector gave up on a partial: wait for every block, then re-arm all slots
        HALA_HIP(hipDeviceSynchronize());
        if (c->red.p) {
            int64_t n = (int64_t)(c->red.bytes / sizeof(double));
            fill_sentinel_kernel<double><<<grid_for(n, 256 * 4, 4096), 256, 0, c->stream>>>(c->red.as<double>(), n);
        }
        if (c->red32.p) {
            int64_t n = (int64_t)(c->red32.bytes / sizeof(float));
            fill_sentinel_kernel<float><<<grid_for(n, 256 * 4, 4096), 256, 0, c->stream>>>(c->red32.as<float>(), n);
        }
        HALA_CHECK_LAUNCH();
    }
    HALA_HIP(hipStreamSynchronize(c->stream));
    if (flag & 4)
        throw status_error{HALA_ERR_NOCONV, "a reduction gave up waiting for a block's partial (device stalled?)"};
    throw status_error{HALA_ERR_NOCONV, "a triangular-solve dependency wait gave up (corrupt schedule?)"};
}

void check_wait_error(hala_ctx c) { check_wait_error_host(c, fetch_i32(c, wait_error_flag(c))); }

}  // namespace hala_amd

hala_ctx hala_amd::aux_ctx(hala_ctx c) {
    if (!c->aux) {
        set_device(c);
        hipStream_t s = nullptr;
        HALA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hala_ctx a = nullptr;
        if (hala_ctx_create(c->device, s, &a) != HALA_OK) {
            (void)hipStreamDestroy(s);
            throw status_error{HALA_ERR_HIP, "auxiliary context: " + std::string(hala_last_error())};
        }
        a->own_stream = true;
        c->aux = a;
    }
    c->aux->opt = c->opt;
    return c->aux;
}

using namespace hala_amd;

extern "C" {

const char *hala_last_error(void) { return g_last_error.c_str(); }

int hala_version(int *major, int *minor) {
    if (major) *major = 1;
    if (minor) *minor = 0;
    return HALA_OK;
}

int hala_device_count(int *count) {
    return guard([&] {
        HALA_REQUIRE(count, HALA_ERR_ARG, "hala_device_count: null count");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = (e == hipSuccess) ? c : 0;
    });
}

int hala_ctx_create(int device, void *stream, hala_ctx *out) {
    return guard([&] {
        HALA_REQUIRE(out, HALA_ERR_ARG, "hala_ctx_create: null output");
        *out = nullptr;
        int count = 0;
        HALA_HIP(hipGetDeviceCount(&count));
        HALA_REQUIRE(device >= 0 && device < count, HALA_ERR_ARG,
                     "hala_ctx_create: device " + std::to_string(device) + " out of range");
        HALA_HIP(hipSetDevice(device));
        auto *c = new hala_ctx_s();
        c->device = device;
        c->stream = (hipStream_t)stream;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;

        try {
            c->scal.reserve(S_TOTAL * sizeof(double));
            HALA_HIP(hipMemsetAsync(c->scal.p, 0, S_TOTAL * sizeof(double), c->stream));
            c->lbc.reserve((size_t)LBC_WORDS * sizeof(unsigned));
            HALA_HIP(hipMemsetAsync(c->lbc.p, 0, c->lbc.bytes, c->stream));
            partials64(c, 64 * 1024);
            partials32(c, 64 * 1024);
            HALA_HIP(hipHostMalloc((void **)&c->host_pinned, 1024 * sizeof(double), hipHostMallocDefault));
            c->noalloc = noalloc_scratch(device);  // the device is current
            HALA_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            if (c->host_pinned) (void)hipHostFree(c->host_pinned);
            delete c;
            throw;
        }
        *out = c;
    });
}

int hala_ctx_destroy(hala_ctx c) {
    return guard([&] {
        if (!c) return;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        if (c->aux) (void)hala_ctx_destroy(c->aux);
        c->tmp.clear();  // temporaries may live in the cache: release them first
        c->cache.reset();
        if (c->host_pinned) (void)hipHostFree(c->host_pinned);
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
    });
}

int hala_ctx_set_stream(hala_ctx c, void *stream) {
    return guard([&] {
        HALA_REQUIRE(c, HALA_ERR_ARG, "null context");
        set_device(c);
        HALA_HIP(hipStreamSynchronize(c->stream));
        c->stream = (hipStream_t)stream;
    });
}

int hala_ctx_sync(hala_ctx c) {
    return guard([&] {
        HALA_REQUIRE(c, HALA_ERR_ARG, "null context");
        set_device(c);
        check_wait_error(c);  // synchronises; reports a wait of earlier asynchronous work that gave up
    });
}

int hala_ctx_device(hala_ctx c, int *device) {
    return guard([&] {
        HALA_REQUIRE(c && device, HALA_ERR_ARG, "null argument");
        *device = c->device;
    });
}

static int *option_slot(hala_ctx c, const char *name) {
    std::string k = name ? name : "";
    if (k == "tri_waves_per_cu") return &c->opt.tri_waves_per_cu;
    if (k == "tri_pool_waves") return &c->opt.tri_pool_waves;
    if (k == "tri_sleep") return &c->opt.tri_sleep;
    if (k == "tri_streams") return &c->opt.tri_streams;
    if (k == "tri_col") return &c->opt.tri_col;
    if (k == "tri_col_cols") return &c->opt.tri_col_cols;
    if (k == "tri_col_split") return &c->opt.tri_col_split;
    if (k == "cg_fuse") return &c->opt.cg_fuse;
    if (k == "cg_xcd") return &c->opt.cg_xcd;
    if (k == "cg_batch_max") return &c->opt.cg_batch_max;
    if (k == "spmv_blocks_per_cu") return &c->opt.spmv_blocks_per_cu;
    if (k == "ilu32_jacobi") return &c->opt.ilu32_jacobi;
    if (k == "ilu64_jacobi") return &c->opt.ilu64_jacobi;
    if (k == "sync_checks") return &c->opt.sync_checks;
    if (k == "tri_stencil") return &c->opt.tri_stencil;
    return nullptr;
}

int hala_ctx_set_option(hala_ctx c, const char *name, int value) {
    return guard([&] {
        HALA_REQUIRE(c, HALA_ERR_ARG, "null context");
        int *slot = option_slot(c, name);
        HALA_REQUIRE(slot, HALA_ERR_ARG, std::string("unknown option ") + (name ? name : "(null)"));
        HALA_REQUIRE(value >= 0 && value <= (1 << 20), HALA_ERR_ARG, "option value out of range");
        HALA_REQUIRE(slot != &c->opt.tri_streams || (value >= 1 && value <= TK_MAX), HALA_ERR_ARG,
                     "tri_streams must lie in 1.." + std::to_string(TK_MAX));
        *slot = value;
    });
}

int hala_ctx_get_option(hala_ctx c, const char *name, int *value) {
    return guard([&] {
        HALA_REQUIRE(c && value, HALA_ERR_ARG, "null argument");
        int *slot = option_slot(c, name);
        HALA_REQUIRE(slot, HALA_ERR_ARG, std::string("unknown option ") + (name ? name : "(null)"));
        *value = *slot;
    });
}

int hala_malloc(hala_ctx c, size_t bytes, void **ptr) {
    return guard([&] {
        HALA_REQUIRE(c && ptr, HALA_ERR_ARG, "hala_malloc: null argument");
        set_device(c);
        *ptr = nullptr;
        if (bytes == 0) return;
        hipError_t e = hipMalloc(ptr, bytes);
        if (e != hipSuccess) {
            *ptr = nullptr;
            throw status_error{HALA_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
        }
    });
}

int hala_free(hala_ctx c, void *ptr) {
    return guard([&] {
        HALA_REQUIRE(c, HALA_ERR_ARG, "hala_free: null context");
        if (!ptr) return;
        set_device(c);
        HALA_HIP(hipStreamSynchronize(c->stream));
        HALA_HIP(hipFree(ptr));
    });
}

int hala_memcpy(hala_ctx c, void *dst, const void *src, size_t bytes, int kind) {
    return guard([&] {
        HALA_REQUIRE(c && kind >= 0 && kind <= 2, HALA_ERR_ARG, "hala_memcpy: bad argument");
        if (bytes == 0) return;
        HALA_REQUIRE(dst && src, HALA_ERR_ARG, "hala_memcpy: null pointer");
        set_device(c);
        hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
        HALA_HIP(hipMemcpyAsync(dst, src, bytes, k, c->stream));
        HALA_HIP(hipStreamSynchronize(c->stream));
    });
}

int hala_fill_f64(hala_ctx c, double *x, int64_t n, double v) { return guard([&] { fill<double>(c, x, n, v); }); }
int hala_fill_f32(hala_ctx c, float *x, int64_t n, float v) { return guard([&] { fill<float>(c, x, n, v); }); }
int hala_fill_i32(hala_ctx c, int *x, int64_t n, int v) { return guard([&] { fill<int>(c, x, n, v); }); }
int hala_convert_f64_f32(hala_ctx c, int64_t n, const double *x, float *y) {
    return guard([&] { convert<double, float>(c, n, x, y); });
}
int hala_convert_f32_f64(hala_ctx c, int64_t n, const float *x, double *y) {
    return guard([&] { convert<float, double>(c, n, x, y); });
}

}  // extern "C"
