// hala.hpp -- drop-in hala:: templates for the sparse iterative-solve path on
// MI355X (gfx950).  The engine, device vector, CSR matrix, triangular matrix,
// ILU(0) and BLAS-1 interfaces mirror the reference's (gpu/hala_gpu_engine.hpp,
// gpu/hala_gpu_vector.hpp, gpu/hala_rocm_sparse_general.hpp,
// gpu/hala_rocm_sparse_triangular.hpp, gpu/hala_gpu_ilu.hpp,
// gpu/hala_gpu_blas1.hpp, gpu/hala_gpu_overloads.hpp); the work happens in
// libhala_amd.so through the C ABI of hala_amd.h, so this header needs only a
// C++14 compiler and the library (no HIP headers).  Library errors become
// std::runtime_error, as in the reference (gpu/hala_rocm_common.hpp:88-109).
//
// Scope: float, double, std::complex<float>, std::complex<double>, unit
// strides, the gpu_engine and the
// mixed_engine (host-resident inputs, hala_mixed.hpp).  Vectors on the device
// are hala::gpu_vector<T> or hala::gpu_wrapped_array<T>.
#pragma once

#include <algorithm>
#include <cassert>
#include <cmath>
#include <complex>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <valarray>
#include <vector>

#include "hala_amd.h"

namespace hala {

inline void check_hala(int status, const char *where) {
    if (status != HALA_OK)
        throw std::runtime_error(std::string("libhala_amd encountered an error with code: ") + std::to_string(status) +
                                 " at " + where + ": " + hala_last_error());
}

// runtime_assert: common/hala_core.hpp:140
inline void runtime_assert(bool condition, const char *message) {
    if (!condition) throw std::runtime_error(message);
}

inline int gpu_device_count() {
    int c = 0;
    check_hala(hala_device_count(&c), "hala_device_count()");
    return c;
}

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
// gpu/hala_gpu_engine.hpp:167-307: device id plus the library context (stream).
// Copies alias the context (non-owning), as the reference's copies alias handles.
class gpu_engine {
  public:
    explicit gpu_engine(int deviceid = 0, void *stream = nullptr) : cgpu(deviceid) {
        hala_ctx c = nullptr;
        check_hala(hala_ctx_create(deviceid, stream, &c), "gpu_engine()");
        ctx = std::shared_ptr<hala_ctx_s>(c, [](hala_ctx p) { hala_ctx_destroy(p); });
    }
    gpu_engine(gpu_engine const &other) = default;
    gpu_engine(gpu_engine &&other) = default;
    gpu_engine &operator=(gpu_engine const &) = default;
    gpu_engine &operator=(gpu_engine &&) = default;

    int device() const { return cgpu; }
    void set_active_device() const {}  // every library call selects the engine's device itself
    void synchronize() const { check_hala(hala_ctx_sync(ctx.get()), "gpu_engine::synchronize()"); }
    void set_stream(void *stream) const { check_hala(hala_ctx_set_stream(ctx.get(), stream), "set_stream()"); }
    hala_ctx handle() const { return ctx.get(); }

    template <class VectorLike>
    auto load(VectorLike const &cpu_data) const;
    template <class VectorLike>
    auto load(VectorLike const &cpu_data, size_t size) const;
    template <class VectorLike>
    auto unload(VectorLike const &gpu_data) const;
    template <typename T>
    auto vector(size_t num_entries, T value) const;

  private:
    int cgpu;
    std::shared_ptr<hala_ctx_s> ctx;
};

// A process-wide engine per device for the device-vector plumbing (allocation,
// copies); deliberately never destroyed, so no library call runs during the
// static destruction at exit.
inline hala_ctx default_ctx(int device) {
    static auto *engines = new std::vector<gpu_engine *>();
    if ((int)engines->size() <= device) engines->resize(device + 1, nullptr);
    if (!(*engines)[device]) (*engines)[device] = new gpu_engine(device);
    return (*engines)[device]->handle();
}

// ---------------------------------------------------------------------------
// host vector adapters (the subset of common/hala_vector_defines.hpp used here)
// ---------------------------------------------------------------------------
template <class V>
struct host_traits;
template <typename T, class A>
struct host_traits<std::vector<T, A>> {
    using value_type = T;
    static T *data(std::vector<T, A> &v) { return v.data(); }
    static T const *data(std::vector<T, A> const &v) { return v.data(); }
    static size_t size(std::vector<T, A> const &v) { return v.size(); }
    static void resize(std::vector<T, A> &v, size_t n) { v.resize(n); }
};
template <typename T>
struct host_traits<std::valarray<T>> {
    using value_type = T;
    static T *data(std::valarray<T> &v) { return std::begin(v); }
    static T const *data(std::valarray<T> const &v) { return std::begin(v); }
    static size_t size(std::valarray<T> const &v) { return v.size(); }
    static void resize(std::valarray<T> &v, size_t n) {
        if (v.size() != n) v.resize(n);
    }
};

// ---------------------------------------------------------------------------
// device vector
// Per-type entry points: real FP64/FP32 and complex<double>/<float> (stored as
// interleaved real, imaginary pairs, as std::complex is).
namespace detail {
inline void fill_(hala_ctx c, double *x, int64_t n, double v) { check_hala(hala_fill_f64(c, x, n, v), "fill()"); }
inline void fill_(hala_ctx c, float *x, int64_t n, float v) { check_hala(hala_fill_f32(c, x, n, v), "fill()"); }
inline void fill_(hala_ctx c, int *x, int64_t n, int v) { check_hala(hala_fill_i32(c, x, n, v), "fill()"); }
template <typename R>
void fill_(hala_ctx c, std::complex<R> *x, int64_t n, std::complex<R> v) {
    if (v.imag() == R(0) && v.real() == R(0)) {
        fill_(c, reinterpret_cast<R *>(x), 2 * n, R(0));
    } else {
        std::vector<std::complex<R>> h((size_t)n, v);
        check_hala(hala_memcpy(c, x, h.data(), (size_t)n * sizeof(std::complex<R>), 0), "fill()");
    }
}
}  // namespace detail

// ---------------------------------------------------------------------------
// gpu/hala_gpu_vector.hpp:49-171: owning device buffer.
template <typename T>
class gpu_vector {
  public:
    using value_type = T;
    explicit gpu_vector(int deviceid = 0) : gpu(deviceid) {}
    gpu_vector(size_t num_entries, int deviceid) : gpu(deviceid) { alloc(num_entries); }
    gpu_vector(gpu_vector const &other) : gpu(other.gpu) {
        alloc(other.num);
        copy_d2d(other.ptr, ptr, num);
    }
    gpu_vector(gpu_vector &&other) noexcept
        : gpu(other.gpu), num(std::exchange(other.num, 0)), ptr(std::exchange(other.ptr, nullptr)) {}
    ~gpu_vector() { clear(); }
    gpu_vector &operator=(gpu_vector const &other) {
        if (this != &other) {
            resize(other.num);
            copy_d2d(other.ptr, ptr, num);
        }
        return *this;
    }
    gpu_vector &operator=(gpu_vector &&other) noexcept {
        std::swap(gpu, other.gpu);
        std::swap(num, other.num);
        std::swap(ptr, other.ptr);
        return *this;
    }
    void clear() {
        if (ptr) hala_free(default_ctx(gpu), ptr);
        ptr = nullptr;
        num = 0;
    }
    void resize(size_t n) {
        if (n == num) return;
        clear();
        alloc(n);
    }
    template <class VectorLike>
    void load(VectorLike const &cpu) {
        using H = host_traits<std::remove_cv_t<VectorLike>>;
        static_assert(std::is_same<T, typename H::value_type>::value, "type mismatch in gpu_vector::load()");
        resize(H::size(cpu));
        if (num)
            check_hala(hala_memcpy(default_ctx(gpu), ptr, H::data(cpu), num * sizeof(T), 0), "gpu_vector::load()");
    }
    template <class VectorLike>
    void unload(VectorLike &cpu) const {
        using H = host_traits<VectorLike>;
        static_assert(std::is_same<T, typename H::value_type>::value, "type mismatch in gpu_vector::unload()");
        H::resize(cpu, num);
        if (num)
            check_hala(hala_memcpy(default_ctx(gpu), H::data(cpu), ptr, num * sizeof(T), 1), "gpu_vector::unload()");
    }
    std::vector<T> unload() const {
        std::vector<T> v(num);
        unload(v);
        return v;
    }
    void fill(T value) {
        if (!num) return;
        detail::fill_(default_ctx(gpu), ptr, (int64_t)num, value);
    }
    T *data() { return ptr; }
    T const *data() const { return ptr; }
    size_t size() const { return num; }
    int device() const { return gpu; }
    bool empty() const { return num == 0; }

  private:
    void alloc(size_t n) {
        num = n;
        ptr = nullptr;
        if (n) {
            void *p = nullptr;
            check_hala(hala_malloc(default_ctx(gpu), n * sizeof(T), &p), "gpu_vector alloc");
            ptr = static_cast<T *>(p);
        }
    }
    void copy_d2d(T const *src, T *dst, size_t n) const {
        if (n) check_hala(hala_memcpy(default_ctx(gpu), dst, src, n * sizeof(T), 2), "gpu_vector copy");
    }
    int gpu = 0;
    size_t num = 0;
    T *ptr = nullptr;
};

// gpu/hala_gpu_wrap_array.hpp:57-80: non-owning view of device memory.
template <typename T>
struct gpu_wrapped_array {
    using value_type = std::remove_cv_t<T>;
    T *ptr;
    size_t num;
    int gpu;
    T *data() const { return ptr; }
    size_t size() const { return num; }
    int device() const { return gpu; }
    void resize(size_t n) { runtime_assert(n <= num, "cannot grow a wrapped device array"); }
};
template <typename T>
gpu_wrapped_array<T> wrap_gpu_array(T *array, size_t num_entries, int device = 0) {
    return {array, num_entries, device};
}

template <class VectorLike>
auto make_gpu_vector(VectorLike const &cpu_data, int gpuid = 0) {
    gpu_vector<typename host_traits<VectorLike>::value_type> v(gpuid);
    v.load(cpu_data);
    return v;
}
template <typename T>
auto make_gpu_vector(size_t num_entries, int gpuid = 0) {
    return gpu_vector<T>(num_entries, gpuid);
}

template <class VectorLike>
auto gpu_engine::load(VectorLike const &cpu_data) const {
    return make_gpu_vector(cpu_data, cgpu);
}
template <class VectorLike>
auto gpu_engine::load(VectorLike const &cpu_data, size_t size) const {
    using T = typename host_traits<VectorLike>::value_type;
    std::vector<T> tmp(host_traits<VectorLike>::data(cpu_data), host_traits<VectorLike>::data(cpu_data) + size);
    return make_gpu_vector(tmp, cgpu);
}
template <class VectorLike>
auto gpu_engine::unload(VectorLike const &gpu_data) const {
    return gpu_data.unload();
}
template <typename T>
auto gpu_engine::vector(size_t num_entries, T value) const {
    gpu_vector<T> x(num_entries, cgpu);
    x.fill(value);
    return x;
}

// binded_gpu_vector (gpu/hala_gpu_vector.hpp:224-249): loads on construction,
// copies back on destruction.
template <typename T, class VectorLike>
struct binded_gpu_vector {
    using value_type = T;
    binded_gpu_vector(int gpuid, VectorLike &v) : vec(v), cvec(gpuid) { cvec.load(v); }
    ~binded_gpu_vector() { cvec.unload(vec); }
    T *data() { return cvec.data(); }
    T const *data() const { return cvec.data(); }
    void resize(size_t n) { cvec.resize(n); }
    size_t size() const { return cvec.size(); }
    int device() const { return cvec.device(); }
    VectorLike &vec;
    gpu_vector<T> cvec;
};
template <class VectorLike>
auto gpu_bind_vector(gpu_engine const &e, VectorLike &x) {
    return binded_gpu_vector<typename host_traits<VectorLike>::value_type, VectorLike>(e.device(), x);
}

// new_vector / set_zero on the engine (gpu/hala_gpu_engine.hpp:327-352)
template <class VectorLike>
auto new_vector(gpu_engine const &engine, VectorLike const &) {
    using T = std::remove_cv_t<typename std::remove_reference_t<VectorLike>::value_type>;
    return gpu_vector<T>(engine.device());
}
template <class VectorLike>
void set_zero(gpu_engine const &engine, size_t num_entries, VectorLike &&x) {
    if (num_entries == 0) return;
    if (x.size() < num_entries) x.resize(num_entries);
    using T = std::remove_cv_t<typename std::remove_reference_t<VectorLike>::value_type>;
    detail::fill_(engine.handle(), reinterpret_cast<T *>(x.data()), (int64_t)num_entries, T(0));
}

template <class V>
using scalar_of = std::remove_cv_t<typename std::remove_reference_t<V>::value_type>;

// Scalar type of a vector and its precision (float for std::complex<float>):
// common/hala_types_checker.hpp:74, :84.  The solvers' stop_criteria are typed
// by the precision, as in the reference.
namespace detail {
template <typename T>
struct precision_of {
    using type = T;
};
template <typename R>
struct precision_of<std::complex<R>> {
    using type = R;
};
}  // namespace detail
template <class V>
using get_standard_type = scalar_of<V>;
template <class V>
using get_precision_type = typename detail::precision_of<scalar_of<V>>::type;

// The vector behind an engine binding, the identity for a plain vector (the
// specialisation for engined_vector is in hala_noengine.hpp).
namespace detail {
template <class V>
struct unbind {
    static V &get(V &v) { return v; }
};
template <class V>
auto &raw(V &v) {
    return unbind<V>::get(v);
}
}  // namespace detail

// Vector sizes (common/hala_vector_defines.hpp:358, :420).
template <class V>
size_t get_size(V const &x) {
    return (size_t)x.size();
}
template <class V>
int get_size_int(V const &x) {
    return static_cast<int>(get_size(x));
}
template <class V>
void set_size(size_t new_size, V &x) {
    x.resize(new_size);
}

// ---------------------------------------------------------------------------
// BLAS-1 on the engine (gpu/hala_gpu_blas1.hpp, gpu/hala_gpu_overloads.hpp)
// ---------------------------------------------------------------------------
namespace detail {
using zd = std::complex<double>;
using zf = std::complex<float>;
inline void copy_(hala_ctx c, int n, double const *x, double *y) { check_hala(hala_copy_f64(c, n, x, y), "vcopy()"); }
inline void copy_(hala_ctx c, int n, float const *x, float *y) { check_hala(hala_copy_f32(c, n, x, y), "vcopy()"); }
inline void copy_(hala_ctx c, int n, zd const *x, zd *y) {
    check_hala(hala_copy_f64(c, 2 * n, reinterpret_cast<double const *>(x), reinterpret_cast<double *>(y)), "vcopy()");
}
inline void copy_(hala_ctx c, int n, zf const *x, zf *y) {
    check_hala(hala_copy_f32(c, 2 * n, reinterpret_cast<float const *>(x), reinterpret_cast<float *>(y)), "vcopy()");
}
inline double nrm2_(hala_ctx c, int64_t n, double const *x) {
    double r;
    check_hala(hala_nrm2_f64(c, n, x, &r), "norm2()");
    return r;
}
inline float nrm2_(hala_ctx c, int64_t n, float const *x) {
    float r;
    check_hala(hala_nrm2_f32(c, n, x, &r), "norm2()");
    return r;
}
inline double nrm2_(hala_ctx c, int64_t n, zd const *x) {
    double r;
    check_hala(hala_nrm2_z(c, n, reinterpret_cast<double const *>(x), &r), "norm2()");
    return r;
}
inline float nrm2_(hala_ctx c, int64_t n, zf const *x) {
    float r;
    check_hala(hala_nrm2_c(c, n, reinterpret_cast<float const *>(x), &r), "norm2()");
    return r;
}
// dot: conjugates x for complex vectors (the reference's dot vs dotu)
inline double dot_(hala_ctx c, int64_t n, double const *x, double const *y, bool) {
    double r;
    check_hala(hala_dot_f64(c, n, x, y, &r), "dot()");
    return r;
}
inline float dot_(hala_ctx c, int64_t n, float const *x, float const *y, bool) {
    float r;
    check_hala(hala_dot_f32(c, n, x, y, &r), "dot()");
    return r;
}
inline zd dot_(hala_ctx c, int64_t n, zd const *x, zd const *y, bool conj) {
    double r[2];
    auto fn = conj ? hala_dotc_z : hala_dotu_z;
    check_hala(fn(c, n, reinterpret_cast<double const *>(x), reinterpret_cast<double const *>(y), r), "dot()");
    return {r[0], r[1]};
}
inline zf dot_(hala_ctx c, int64_t n, zf const *x, zf const *y, bool conj) {
    float r[2];
    auto fn = conj ? hala_dotc_c : hala_dotu_c;
    check_hala(fn(c, n, reinterpret_cast<float const *>(x), reinterpret_cast<float const *>(y), r), "dot()");
    return {r[0], r[1]};
}
inline void axpy_(hala_ctx c, int64_t n, double a, double const *x, double *y) {
    check_hala(hala_axpy_f64(c, n, a, x, y), "axpy()");
}
inline void axpy_(hala_ctx c, int64_t n, float a, float const *x, float *y) {
    check_hala(hala_axpy_f32(c, n, a, x, y), "axpy()");
}
inline void axpy_(hala_ctx c, int64_t n, zd a, zd const *x, zd *y) {
    check_hala(hala_axpy_z(c, n, a.real(), a.imag(), reinterpret_cast<double const *>(x), reinterpret_cast<double *>(y)),
               "axpy()");
}
inline void axpy_(hala_ctx c, int64_t n, zf a, zf const *x, zf *y) {
    check_hala(hala_axpy_c(c, n, a.real(), a.imag(), reinterpret_cast<float const *>(x), reinterpret_cast<float *>(y)),
               "axpy()");
}
inline void scal_(hala_ctx c, int64_t n, double a, double *x) { check_hala(hala_scal_f64(c, n, a, x), "scal()"); }
inline void scal_(hala_ctx c, int64_t n, float a, float *x) { check_hala(hala_scal_f32(c, n, a, x), "scal()"); }
inline void scal_(hala_ctx c, int64_t n, zd a, zd *x) {
    check_hala(hala_scal_z(c, n, a.real(), a.imag(), reinterpret_cast<double *>(x)), "scal()");
}
inline void scal_(hala_ctx c, int64_t n, zf a, zf *x) {
    check_hala(hala_scal_c(c, n, a.real(), a.imag(), reinterpret_cast<float *>(x)), "scal()");
}
inline void spmv_(hala_csr A, char t, double a, double const *v, double const *x, double b, double *y) {
    check_hala(hala_spmv_f64(A, t, a, v, x, b, y), "gpu_sparse_matrix::gemv()");
}
inline void spmv_(hala_csr A, char t, float a, float const *v, float const *x, float b, float *y) {
    check_hala(hala_spmv_f32(A, t, a, v, x, b, y), "gpu_sparse_matrix::gemv()");
}
inline void spmv_(hala_csr A, char t, zd a, zd const *v, zd const *x, zd b, zd *y) {
    check_hala(hala_spmv_z(A, t, a.real(), a.imag(), reinterpret_cast<double const *>(v),
                           reinterpret_cast<double const *>(x), b.real(), b.imag(), reinterpret_cast<double *>(y)),
               "gpu_sparse_matrix::gemv()");
}
inline void spmv_(hala_csr A, char t, zf a, zf const *v, zf const *x, zf b, zf *y) {
    check_hala(hala_spmv_c(A, t, a.real(), a.imag(), reinterpret_cast<float const *>(v),
                           reinterpret_cast<float const *>(x), b.real(), b.imag(), reinterpret_cast<float *>(y)),
               "gpu_sparse_matrix::gemv()");
}
inline void spmm_(hala_csr A, char ta, char tb, int br, int bc, double a, double const *v, double const *B, int ldb,
                  double b, double *C, int ldc) {
    check_hala(hala_spmm_f64(A, ta, tb, br, bc, a, v, B, ldb, b, C, ldc), "gpu_sparse_matrix::gemm()");
}
inline void spmm_(hala_csr A, char ta, char tb, int br, int bc, float a, float const *v, float const *B, int ldb,
                  float b, float *C, int ldc) {
    check_hala(hala_spmm_f32(A, ta, tb, br, bc, a, v, B, ldb, b, C, ldc), "gpu_sparse_matrix::gemm()");
}
inline void spmm_(hala_csr A, char ta, char tb, int br, int bc, zd a, zd const *v, zd const *B, int ldb, zd b, zd *C,
                  int ldc) {
    check_hala(hala_spmm_z(A, ta, tb, br, bc, a.real(), a.imag(), reinterpret_cast<double const *>(v),
                           reinterpret_cast<double const *>(B), ldb, b.real(), b.imag(), reinterpret_cast<double *>(C),
                           ldc),
               "gpu_sparse_matrix::gemm()");
}
inline void spmm_(hala_csr A, char ta, char tb, int br, int bc, zf a, zf const *v, zf const *B, int ldb, zf b, zf *C,
                  int ldc) {
    check_hala(hala_spmm_c(A, ta, tb, br, bc, a.real(), a.imag(), reinterpret_cast<float const *>(v),
                           reinterpret_cast<float const *>(B), ldb, b.real(), b.imag(), reinterpret_cast<float *>(C),
                           ldc),
               "gpu_sparse_matrix::gemm()");
}
}  // namespace detail

template <class VX, class VY>
void vcopy(gpu_engine const &engine, int N, VX const &x, int incx, VY &&y, int incy) {
    runtime_assert(incx == 1 && incy == 1, "hala_amd: unit strides only");
    if ((int)y.size() < N) y.resize(N);
    using T = scalar_of<VX>;
    detail::copy_(engine.handle(), N, (T const *)x.data(), (T *)y.data());
}
template <class VX, class VY>
void vcopy(gpu_engine const &engine, VX const &x, VY &&y) {
    vcopy(engine, (int)x.size(), x, 1, y, 1);
}
// norm2: the precision type (double for complex<double>)
template <class VX>
auto norm2(gpu_engine const &engine, VX const &x) {
    using T = scalar_of<VX>;
    return detail::nrm2_(engine.handle(), (int64_t)x.size(), (T const *)x.data());
}
template <class VX, class VY>
auto dot(gpu_engine const &engine, VX const &x, VY const &y) {
    using T = scalar_of<VX>;
    return detail::dot_(engine.handle(), (int64_t)x.size(), (T const *)x.data(), (T const *)y.data(), true);
}
template <class VX, class VY>
auto dotu(gpu_engine const &engine, VX const &x, VY const &y) {
    using T = scalar_of<VX>;
    return detail::dot_(engine.handle(), (int64_t)x.size(), (T const *)x.data(), (T const *)y.data(), false);
}
template <typename FP, class VX, class VY>
void axpy(gpu_engine const &engine, FP alpha, VX const &x, VY &&y) {
    using T = scalar_of<VX>;
    detail::axpy_(engine.handle(), (int64_t)x.size(), (T)alpha, (T const *)x.data(), (T *)y.data());
}
template <typename FP, class VX>
void scal(gpu_engine const &engine, FP alpha, VX &&x) {
    using T = scalar_of<VX>;
    detail::scal_(engine.handle(), (int64_t)x.size(), (T)alpha, (T *)x.data());
}
// dense column-major gemv (gpu/hala_gpu_blas2.hpp:40-60), FP64
template <class VA, class VX, class VY>
void gemv(gpu_engine const &engine, char trans, int M, int N, double alpha, VA const &A, VX const &x, double beta,
          VY &&y) {
    int need = (trans == 'N' || trans == 'n') ? M : N;
    if ((int)y.size() < need) y.resize(need);
    check_hala(hala_gemv_f64(engine.handle(), trans, M, N, alpha, (double const *)A.data(), M > 1 ? M : 1,
                             (double const *)x.data(), beta, (double *)y.data()),
               "gemv()");
}

// ---------------------------------------------------------------------------
// CSR matrix (gpu/hala_rocm_sparse_general.hpp:74-199)
// ---------------------------------------------------------------------------
template <typename T>
class gpu_sparse_matrix {
  public:
    using value_type = T;
    using engine_type = gpu_engine;
    template <class VP, class VI, class VV>
    gpu_sparse_matrix(gpu_engine const &engine, int num_rows, int num_cols, int num_nz, VP const &pntr, VI const &indx,
                      VV const &vals)
        : rengine(engine), rvals(vals.data()), rows(num_rows), cols(num_cols), nnz(num_nz) {
        hala_csr h = nullptr;
        check_hala(hala_csr_create(engine.handle(), num_rows, num_cols, num_nz, pntr.data(), indx.data(), &h),
                   "make_sparse_matrix()");
        csr = std::shared_ptr<hala_csr_s>(h, [](hala_csr p) { hala_csr_destroy(p); });
    }
    template <class VP, class VI, class VV>
    gpu_sparse_matrix(gpu_engine const &engine, int num_cols, VP const &pntr, VI const &indx, VV const &vals)
        : gpu_sparse_matrix(engine, (int)pntr.size() - 1, num_cols, (int)indx.size(), pntr, indx, vals) {}

    gpu_engine const &engine() const { return rengine; }
    hala_csr handle() const { return csr.get(); }
    int num_rows() const { return rows; }

    template <typename FPa, class VX, typename FPb, class VY>
    size_t gemv_buffer_size(char, FPa, VX const &, FPb, VY &&) const {
        return 0;
    }
    template <typename FPa, class VX, typename FPb, class VY, class VB>
    void gemv(char trans, FPa alpha, VX const &x, FPb beta, VY &&y, VB &&) const {
        gemv(trans, alpha, x, beta, y);
    }
    template <typename FPa, class VX, typename FPb, class VY>
    void gemv(char trans, FPa alpha, VX const &x, FPb beta, VY &&y) const {
        size_t out = (trans == 'N' || trans == 'n') ? rows : cols;
        if (y.size() < out) {
            runtime_assert(beta == FPb(0), "gemv: y too small for beta != 0");
            y.resize(out);
        }
        runtime_assert(beta == FPb(0) || y.size() >= out, "gemv: y too small");
        detail::spmv_(csr.get(), trans, (T)alpha, rvals, (T const *)x.data(), (T)beta, (T *)y.data());
    }
    // C = alpha op(A) op(B) + beta C (gpu/hala_rocm_sparse_general.hpp:131-161); C grows to
    // ldc x N when beta == 0 and it is smaller
    template <typename FSA, class VB, typename FSB, class VC>
    size_t gemm_buffer_size(char, char, int, int, FSA, VB const &, int, FSB, VC &&, int) const {
        return 0;
    }
    template <typename FSA, class VB, typename FSB, class VC, class VT>
    void gemm(char transa, char transb, int b_rows, int b_cols, FSA alpha, VB const &B, int ldb, FSB beta, VC &&C,
              int ldc, VT &&) const {
        gemm(transa, transb, b_rows, b_cols, alpha, B, ldb, beta, C, ldc);
    }
    template <typename FSA, class VB, typename FSB, class VC>
    void gemm(char transa, char transb, int b_rows, int b_cols, FSA alpha, VB const &B, int ldb, FSB beta, VC &&C,
              int ldc) const {
        const int N = (transb == 'N' || transb == 'n') ? b_cols : b_rows;
        if ((size_t)C.size() < (size_t)ldc * (size_t)N) {
            runtime_assert(beta == FSB(0), "gemm: C too small for beta != 0");
            C.resize((size_t)ldc * (size_t)N);
        }
        detail::spmm_(csr.get(), transa, transb, b_rows, b_cols, (T)alpha, rvals, (T const *)B.data(), ldb, (T)beta,
                      (T *)C.data(), ldc);
    }
    T const *values() const { return rvals; }

  private:
    gpu_engine rengine;
    T const *rvals;
    int rows, cols, nnz;
    std::shared_ptr<hala_csr_s> csr;
};

template <class VP, class VI, class VV>
auto make_sparse_matrix(gpu_engine const &engine, int num_rows, int num_cols, int num_nz, VP const &pntr,
                        VI const &indx, VV const &vals) {
    return gpu_sparse_matrix<scalar_of<VV>>(engine, num_rows, num_cols, num_nz, pntr, indx, vals);
}
template <class VP, class VI, class VV>
auto make_sparse_matrix(gpu_engine const &engine, int num_cols, VP const &pntr, VI const &indx, VV const &vals) {
    return gpu_sparse_matrix<scalar_of<VV>>(engine, num_cols, pntr, indx, vals);
}
// sparse/hala_sparse_blas.hpp:217-224
template <typename MT, typename FSA, class VB, typename FSB, class VC>
void sparse_gemm(MT const &matrix, char transa, char transb, int b_rows, int b_cols, FSA alpha, VB const &B, int ldb,
                 FSB beta, VC &&C, int ldc) {
    matrix.gemm(transa, transb, b_rows, b_cols, alpha, detail::raw(B), ldb, beta, detail::raw(C), ldc);
}
template <typename FPa, class VP, class VI, class VV, class VX, typename FPb, class VY>
void sparse_gemv(gpu_engine const &engine, char trans, int M, int N, FPa alpha, VP const &pntr, VI const &indx,
                 VV const &vals, VX const &x, FPb beta, VY &&y) {
    make_sparse_matrix(engine, M, N, (int)indx.size(), pntr, indx, vals).gemv(trans, alpha, x, beta, y);
}
template <class M, typename FPa, class VX, typename FPb, class VY, class VB>
void sparse_gemv(M const &matrix, char trans, FPa alpha, VX const &x, FPb beta, VY &&y, VB &&buf) {
    matrix.gemv(trans, alpha, detail::raw(x), beta, detail::raw(y), buf);
}
template <class M, typename FPa, class VX, typename FPb, class VY>
void sparse_gemv(M const &matrix, char trans, FPa alpha, VX const &x, FPb beta, VY &&y) {
    matrix.gemv(trans, alpha, detail::raw(x), beta, detail::raw(y));
}
template <class M, typename FPa, class VX, typename FPb, class VY>
size_t sparse_gemv_buffer_size(M const &matrix, char trans, FPa alpha, VX const &x, FPb beta, VY &&y) {
    return matrix.gemv_buffer_size(trans, alpha, detail::raw(x), beta, detail::raw(y));
}

// ---------------------------------------------------------------------------
// triangular matrix (gpu/hala_rocm_sparse_triangular.hpp:29-194)
// ---------------------------------------------------------------------------
template <typename T>
class gpu_triangular_matrix {
  public:
    using value_type = T;
    using engine_type = gpu_engine;
    template <class VP, class VI, class VV>
    gpu_triangular_matrix(gpu_engine const &engine, char uplo, char diag, VP const &pntr, VI const &indx,
                          VV const &vals, char /*policy*/ = 'N')
        : rengine(engine), rvals(vals.data()), n((int)pntr.size() - 1) {
        hala_tri h = nullptr;
        check_hala(hala_tri_create(engine.handle(), uplo, diag, n, (int64_t)indx.size(), pntr.data(), indx.data(), &h),
                   "make_triangular_matrix()");
        tri = std::shared_ptr<hala_tri_s>(h, [](hala_tri p) { hala_tri_destroy(p); });
    }
    template <typename FP, class VB, class VX>
    size_t trsv_buffer_size(char, FP, VB const &, VX &&) const {
        return 0;
    }
    template <typename FP, class VB, class VX, class VW>
    void trsv(char trans, FP alpha, VB const &b, VX &&x, VW &&) const {
        trsv(trans, alpha, b, x);
    }
    template <typename FP, class VB, class VX>
    void trsv(char trans, FP alpha, VB const &b, VX &&x) const {
        if ((int)x.size() < n) x.resize(n);
        check_hala(trsv_(trans, T(alpha), b.data(), x.data()), "gpu_triangular_matrix::trsv()");
    }
    // B <- op(T)^-1 (alpha B) column by column (sparse/hala_sparse_structs.hpp:278-300)
    template <typename FP, class VB>
    size_t trsm_buffer_size(char, char, int, FP, VB &&, int = -1) const {
        return 0;
    }
    template <typename FP, class VB, class VW>
    void trsm(char transa, char transb, int nrhs, FP alpha, VB &&B, int ldb, VW &&) const {
        trsm(transa, transb, nrhs, alpha, B, ldb);
    }
    template <typename FP, class VB>
    void trsm(char transa, char transb, int nrhs, FP alpha, VB &&B, int ldb = -1) const {
        check_hala(trsm_(transa, transb, nrhs, T(alpha), B.data(), ldb), "gpu_triangular_matrix::trsm()");
    }

  private:
    int trsv_(char t, double a, double const *b, double *x) const {
        return hala_trsv_f64(tri.get(), t, a, (double const *)rvals, b, x);
    }
    int trsv_(char t, float a, float const *b, float *x) const {
        return hala_trsv_f32(tri.get(), t, a, (float const *)rvals, b, x);
    }
    int trsv_(char t, std::complex<double> a, std::complex<double> const *b, std::complex<double> *x) const {
        return hala_trsv_z(tri.get(), t, a.real(), a.imag(), (double const *)rvals, (double const *)b, (double *)x);
    }
    int trsv_(char t, std::complex<float> a, std::complex<float> const *b, std::complex<float> *x) const {
        return hala_trsv_c(tri.get(), t, a.real(), a.imag(), (float const *)rvals, (float const *)b, (float *)x);
    }
    int trsm_(char ta, char tb, int k, double a, double *B, int ldb) const {
        return hala_trsm_f64(tri.get(), ta, tb, k, a, (double const *)rvals, B, ldb);
    }
    int trsm_(char ta, char tb, int k, float a, float *B, int ldb) const {
        return hala_trsm_f32(tri.get(), ta, tb, k, a, (float const *)rvals, B, ldb);
    }
    int trsm_(char ta, char tb, int k, std::complex<double> a, std::complex<double> *B, int ldb) const {
        return hala_trsm_z(tri.get(), ta, tb, k, a.real(), a.imag(), (double const *)rvals, (double *)B, ldb);
    }
    int trsm_(char ta, char tb, int k, std::complex<float> a, std::complex<float> *B, int ldb) const {
        return hala_trsm_c(tri.get(), ta, tb, k, a.real(), a.imag(), (float const *)rvals, (float *)B, ldb);
    }
    gpu_engine rengine;
    T const *rvals;
    int n;
    std::shared_ptr<hala_tri_s> tri;
};
template <class VP, class VI, class VV>
auto make_triangular_matrix(gpu_engine const &engine, char uplo, char diag, VP const &pntr, VI const &indx,
                            VV const &vals, char policy = 'N') {
    return gpu_triangular_matrix<scalar_of<VV>>(engine, uplo, diag, pntr, indx, vals, policy);
}
template <class TM, typename FP, class VB, class VX>
void sparse_trsv(char trans, TM const &tri, FP alpha, VB const &b, VX &&x) {
    tri.trsv(trans, alpha, detail::raw(b), detail::raw(x));
}
// sparse/hala_sparse_blas.hpp:203-205
template <class TM, typename FP, class VB>
void sparse_trsm(char transa, char transb, int nrhs, TM const &tri, FP alpha, VB &&B, int ldb = -1) {
    tri.trsm(transa, transb, nrhs, alpha, detail::raw(B), ldb);
}

// ---------------------------------------------------------------------------
// ILU(0) (gpu/hala_gpu_ilu.hpp:45-195, wax/hala_sparse_extensions.hpp:51-107)
// ---------------------------------------------------------------------------
template <typename T>
class gpu_ilu {
  public:
    using value_type = T;
    using engine_type = gpu_engine;
    template <class VP, class VI, class VV>
    gpu_ilu(gpu_engine const &engine, VP const &pntr, VI const &indx, VV const &vals, char policy = 'N',
            bool want_f32 = false)
        : rengine(engine), n((int)pntr.size() - 1), rpolicy(policy) {
        hala_ilu h = nullptr;
        check_hala(hala_ilu_create(engine.handle(), n, (int64_t)indx.size(), pntr.data(), indx.data(), &h),
                   "make_ilu()");
        ilu = std::shared_ptr<hala_ilu_s>(h, [](hala_ilu p) { hala_ilu_destroy(p); });
        factor(engine, (T const *)vals.data(), (int64_t)indx.size(), want_f32);
    }
    template <class VX, class VR>
    size_t buffer_size(VX const &, VR &&, int) const {
        return 0;
    }
    template <class VX, class VR>
    gpu_vector<T> get_temp_buffer(VX const &, VR &&, int) const {
        return gpu_vector<T>(rengine.device());
    }
    template <class VX, class VR, class VT>
    void apply(VX const &x, VR &&r, int num_rhs, VT &&) const {
        apply(x, r, num_rhs);
    }
    template <class VX, class VR>
    void apply(VX const &x, VR &&r, int num_rhs = 1) const {
        // num_rhs columns of n entries, column-major (gpu/hala_gpu_ilu.hpp:140-195)
        runtime_assert(num_rhs >= 1 && x.size() >= (size_t)n * num_rhs, "gpu_ilu::apply(): x too small");
        if (r.size() < (size_t)n * num_rhs) r.resize((size_t)n * num_rhs);
        for (int k = 0; k < num_rhs; k++)
            apply_(ilu.get(), (T const *)x.data() + (size_t)k * n, (T *)r.data() + (size_t)k * n);
    }
    gpu_engine const &engine() const { return rengine; }
    hala_ilu handle() const { return ilu.get(); }

  private:
    // FP64 factors (FP32 copies on request); FP32 values are factored in FP64 and applied in FP32;
    // complex values in their own precision
    void factor(gpu_engine const &, double const *v, int64_t, bool want_f32) {
        check_hala(hala_ilu_factor_f64(ilu.get(), v, want_f32 ? 1 : 0), "make_ilu()");
    }
    void factor(gpu_engine const &engine, float const *v, int64_t nnz, bool) {
        gpu_vector<double> v64((size_t)nnz, engine.device());
        check_hala(hala_convert_f32_f64(engine.handle(), nnz, v, v64.data()), "make_ilu()");
        check_hala(hala_ilu_factor_f64(ilu.get(), v64.data(), 1), "make_ilu()");
    }
    void factor(gpu_engine const &, std::complex<double> const *v, int64_t, bool) {
        check_hala(hala_ilu_factor_z(ilu.get(), reinterpret_cast<double const *>(v)), "make_ilu()");
    }
    void factor(gpu_engine const &, std::complex<float> const *v, int64_t, bool) {
        check_hala(hala_ilu_factor_c(ilu.get(), reinterpret_cast<float const *>(v)), "make_ilu()");
    }
    static void apply_(hala_ilu h, double const *x, double *r) {
        check_hala(hala_ilu_apply_f64(h, x, r), "gpu_ilu::apply()");
    }
    static void apply_(hala_ilu h, float const *x, float *r) {
        check_hala(hala_ilu_apply_f32(h, x, r), "gpu_ilu::apply()");
    }
    static void apply_(hala_ilu h, std::complex<double> const *x, std::complex<double> *r) {
        check_hala(hala_ilu_apply_z(h, reinterpret_cast<double const *>(x), reinterpret_cast<double *>(r)),
                   "gpu_ilu::apply()");
    }
    static void apply_(hala_ilu h, std::complex<float> const *x, std::complex<float> *r) {
        check_hala(hala_ilu_apply_c(h, reinterpret_cast<float const *>(x), reinterpret_cast<float *>(r)),
                   "gpu_ilu::apply()");
    }
    gpu_engine rengine;
    int n;
    char rpolicy;
    std::shared_ptr<hala_ilu_s> ilu;
};

template <class VP, class VI, class VV>
auto make_ilu(gpu_engine const &engine, VP const &pntr, VI const &indx, VV const &vals, char policy = 'N') {
    return gpu_ilu<scalar_of<VV>>(engine, pntr, indx, vals, policy);
}
template <class ILU, class VX, class VR, class VT>
void ilu_apply(ILU const &ilu, VX const &x, VR &&r, int num_rhs, VT &&tmp) {
    ilu.apply(detail::raw(x), detail::raw(r), num_rhs, tmp);
}
template <class ILU, class VX, class VR>
void ilu_apply(ILU const &ilu, VX const &x, VR &&r, int num_rhs = 1) {
    ilu.apply(detail::raw(x), detail::raw(r), num_rhs);
}
template <class ILU, class VX, class VR>
auto ilu_temp_buffer(ILU const &ilu, VX const &x, VR &&r, int num_rhs) {
    return ilu.get_temp_buffer(detail::raw(x), detail::raw(r), num_rhs);
}
template <class ILU, class VX, class VR>
size_t ilu_buffer_size(ILU const &ilu, VX const &x, VR &&r, int num_rhs) {
    return ilu.buffer_size(detail::raw(x), detail::raw(r), num_rhs);
}

}  // namespace hala

#include "hala_mixed.hpp"
#include "hala_noengine.hpp"
