// blas1.hip -- BLAS-1 and dense gemv on the engine (the rocBLAS calls of
// gpu/hala_gpu_blas1.hpp and gpu/hala_gpu_blas2.hpp:40-60 in the reference).
#include "blas1.hpp"
#include "common.hpp"

namespace hala_amd {

template <class T, int K>
__global__ __launch_bounds__(256) void cdot_kernel(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                   int64_t nb, T *part, int *err, T *out, int sq) {
    __shared__ T sh[CB * K];
    T v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        int64_t i = ((int64_t)blockIdx.x * K + k) * CB + threadIdx.x;
        T a = T(0);
        if (i < n) a = sq ? x[i] * x[i] : x[i] * y[i];
        v[k] = a;
    }
    tree256<T, K>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            int64_t b = (int64_t)blockIdx.x * K + k;
            if (b < nb) st_agent(part + b, v[k]);
        }
    }
    if (is_collector()) {
        T r = final_pass<T>(part, nb, sh, err);
        if (threadIdx.x == 0) out[0] = sq == 2 ? (T)sqrt(r) : r;
    }
}

template <class T>
void cdot_launch(hala_ctx c, int64_t n, const T *x, const T *y, T *out_dev, int mode) {
    constexpr int K = 4;
    int64_t nb = (n + CB - 1) / CB;
    if (nb < 1) nb = 1;
    T *part;
    if constexpr (sizeof(T) == 8) part = reinterpret_cast<T *>(partials64(c, nb));
    else part = reinterpret_cast<T *>(partials32(c, nb));
    int64_t g = (nb + K - 1) / K;
    HALA_REQUIRE(g < (1ll << 31), HALA_ERR_ARG, "vector too long");
    cdot_kernel<T, K><<<(unsigned)g, CB, 0, c->stream>>>(n, x, y, nb, part, wait_error_flag(c), out_dev, mode);
    HALA_CHECK_LAUNCH();
}
template void cdot_launch<double>(hala_ctx, int64_t, const double *, const double *, double *, int);
template void cdot_launch<float>(hala_ctx, int64_t, const float *, const float *, float *, int);

template <class T>
__global__ void axpy_kernel(int64_t n, T a, const T *__restrict__ x, T *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) y[i] = y[i] + a * x[i];
}
template <class T>
__global__ void scal_kernel(int64_t n, T a, T *__restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) x[i] = a * x[i];
}
template <class T>
__global__ void copy_kernel(int64_t n, const T *__restrict__ x, T *__restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) y[i] = x[i];
}
__global__ void vdivide_kernel(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                               double *__restrict__ z) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) z[i] = x[i] / y[i];
}

// y = beta*y (beta==0: no read) then y_i = y_i + (alpha*x_j)*A_ij for j ascending:
// the element order of reference dgemv 'N'.
__global__ void gemv_n_kernel(int m, int n, double alpha, const double *__restrict__ A, int lda,
                              const double *__restrict__ x, double beta, double *__restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double yi = (beta == 0.0) ? 0.0 : ((beta == 1.0) ? y[i] : beta * y[i]);
    for (int j = 0; j < n; j++) yi = yi + (alpha * x[j]) * A[(int64_t)j * lda + i];
    y[i] = yi;
}

// y_j = alpha * cdot(A_j, x) + beta*y_j, one column per blockIdx.y; partials per column.
template <int K>
__global__ __launch_bounds__(256) void gemv_t_kernel(int m, const double *__restrict__ A, int lda,
                                                     const double *__restrict__ x, double alpha, double beta,
                                                     double *__restrict__ y, int64_t nb, double *part,
                                                     int *err) {
    __shared__ double sh[CB * K];
    const int col = blockIdx.y;
    const double *a = A + (int64_t)col * lda;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        int64_t i = ((int64_t)blockIdx.x * K + k) * CB + threadIdx.x;
        v[k] = (i < m) ? a[i] * x[i] : 0.0;
    }
    tree256<double, K>(v, sh);
    double *pc = part + (int64_t)col * nb;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            int64_t b = (int64_t)blockIdx.x * K + k;
            if (b < nb) st_agent(pc + b, v[k]);
        }
    }
    if (is_collector()) {
        double r = final_pass<double>(pc, nb, sh, err);
        if (threadIdx.x == 0) y[col] = (beta == 0.0) ? alpha * r + 0.0 : alpha * r + beta * y[col];
    }
}

template <class T>
static T dot_host(hala_ctx c, int64_t n, const T *x, const T *y, int mode) {
    HALA_REQUIRE(c && n >= 0, HALA_ERR_ARG, "dot: bad argument");
    HALA_REQUIRE(n == 0 || (x && (y || mode)), HALA_ERR_ARG, "dot: null vector");
    set_device(c);
    if (n == 0) return T(0);
    T *out = reinterpret_cast<T *>(c->scal.as<double>() + S_DOT);
    cdot_launch<T>(c, n, x, y, out, mode);
    // one copy brings back the result and the wait-error flag (same scalar block)
    const char *lo = reinterpret_cast<const char *>(out);
    const size_t span = reinterpret_cast<const char *>(wait_error_flag(c) + 1) - lo;
    HALA_HIP(hipMemcpyAsync(c->host_pinned, lo, span, hipMemcpyDeviceToHost, c->stream));
    HALA_HIP(hipStreamSynchronize(c->stream));
    int flag;
    std::memcpy(&flag, reinterpret_cast<const char *>(c->host_pinned) + (span - sizeof(int)), sizeof(int));
    const T r = *reinterpret_cast<T *>(c->host_pinned);
    check_wait_error_host(c, flag);
    return r;
}

template <class K_, class... A>
static void launch_1d(hala_ctx c, int64_t n, K_ kern, A... args) {
    if (n <= 0) return;
    int g = grid_for(n, 256 * 4, 16384);
    kern<<<g, 256, 0, c->stream>>>(n, args...);
    HALA_CHECK_LAUNCH();
}

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

int hala_dot_f64(hala_ctx c, int64_t n, const double *x, const double *y, double *res) {
    return guard([&] {
        HALA_REQUIRE(res, HALA_ERR_ARG, "dot: null result");
        *res = dot_host<double>(c, n, x, y, 0);
    });
}
int hala_nrm2_f64(hala_ctx c, int64_t n, const double *x, double *res) {
    return guard([&] {
        HALA_REQUIRE(res, HALA_ERR_ARG, "nrm2: null result");
        *res = dot_host<double>(c, n, x, x, 2);
    });
}
int hala_dot_f32(hala_ctx c, int64_t n, const float *x, const float *y, float *res) {
    return guard([&] {
        HALA_REQUIRE(res, HALA_ERR_ARG, "dot: null result");
        *res = dot_host<float>(c, n, x, y, 0);
    });
}
int hala_nrm2_f32(hala_ctx c, int64_t n, const float *x, float *res) {
    return guard([&] {
        HALA_REQUIRE(res, HALA_ERR_ARG, "nrm2: null result");
        *res = dot_host<float>(c, n, x, x, 2);
    });
}

int hala_axpy_f64(hala_ctx c, int64_t n, double a, const double *x, double *y) {
    return guard([&] {
        HALA_REQUIRE(c && n >= 0 && (n == 0 || (x && y)), HALA_ERR_ARG, "axpy: bad argument");
        set_device(c);
        launch_1d(c, n, axpy_kernel<double>, a, x, y);
    });
}
int hala_axpy_f32(hala_ctx c, int64_t n, float a, const float *x, float *y) {
    return guard([&] {
        HALA_REQUIRE(c && n >= 0 && (n == 0 || (x && y)), HALA_ERR_ARG, "axpy: bad argument");
        set_device(c);
        launch_1d(c, n, axpy_kernel<float>, a, x, y);
    });
}
int hala_scal_f64(hala_ctx c, int64_t n, double a, double *x) {
    return guard([&] {
        HALA_REQUIRE(c && n >= 0 && (n == 0 || x), HALA_ERR_ARG, "scal: bad argument");
        set_device(c);
        launch_1d(c, n, scal_kernel<double>, a, x);
    });
}
int hala_scal_f32(hala_ctx c, int64_t n, float a, float *x) {
    return guard([&] {
        HALA_REQUIRE(c && n >= 0 && (n == 0 || x), HALA_ERR_ARG, "scal: bad argument");
        set_device(c);
        launch_1d(c, n, scal_kernel<float>, a, x);
    });
}
int hala_copy_f64(hala_ctx c, int64_t n, const double *x, double *y) {
    return guard([&] {
        HALA_REQUIRE(c && n >= 0 && (n == 0 || (x && y)), HALA_ERR_ARG, "copy: bad argument");
        set_device(c);
        launch_1d(c, n, copy_kernel<double>, x, y);
    });
}
int hala_copy_f32(hala_ctx c, int64_t n, const float *x, float *y) {
    return guard([&] {
        HALA_REQUIRE(c && n >= 0 && (n == 0 || (x && y)), HALA_ERR_ARG, "copy: bad argument");
        set_device(c);
        launch_1d(c, n, copy_kernel<float>, x, y);
    });
}
int hala_vdivide_f64(hala_ctx c, int64_t n, const double *x, const double *y, double *z) {
    return guard([&] {
        HALA_REQUIRE(c && n >= 0 && (n == 0 || (x && y && z)), HALA_ERR_ARG, "vdivide: bad argument");
        set_device(c);
        launch_1d(c, n, vdivide_kernel, x, y, z);
    });
}

int hala_gemv_f64(hala_ctx c, char trans, int m, int n, double alpha, const double *A, int lda, const double *x,
                  double beta, double *y) {
    return guard([&] {
        HALA_REQUIRE(c && m >= 0 && n >= 0 && lda >= (m > 1 ? m : 1), HALA_ERR_ARG, "gemv: bad dimensions");
        set_device(c);
        if (trans == 'N' || trans == 'n') {
            if (m == 0) return;
            HALA_REQUIRE(A && x && y, HALA_ERR_ARG, "gemv: null pointer");
            gemv_n_kernel<<<(m + 255) / 256, 256, 0, c->stream>>>(m, n, alpha, A, lda, x, beta, y);
            HALA_CHECK_LAUNCH();
        } else {
            if (n == 0) return;
            HALA_REQUIRE(A && x && y, HALA_ERR_ARG, "gemv: null pointer");
            HALA_REQUIRE(n <= 32, HALA_ERR_ARG, "gemv 'T': at most 32 columns (the Krylov basis width)");
            constexpr int K = 4;
            int64_t nb = (m + CB - 1) / CB;
            if (nb < 1) nb = 1;
            double *part = partials64(c, (size_t)nb * n);
            dim3 g((unsigned)((nb + K - 1) / K), (unsigned)n);
            gemv_t_kernel<K><<<g, CB, 0, c->stream>>>(m, A, lda, x, alpha, beta, y, nb, part,
                                                     wait_error_flag(c));
            HALA_CHECK_LAUNCH();
        }
    });
}

}  // extern "C"
