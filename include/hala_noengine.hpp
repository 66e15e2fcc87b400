// hala_noengine.hpp -- the no-engine form of the API: a vector bound to an
// engine carries that engine, and calls without an engine argument take it
// from their bound vectors (wax/hala_noengine_structs.hpp:154-311 engined_vector
// and bind_engine_vector, wax/hala_noengine_overloads.hpp the BLAS / sparse
// overloads).  User code written once against bound vectors -- the style of
// examples/example_gpu.cpp:208-285 -- then runs on the gpu_engine (device
// vectors) or the mixed_engine (host vectors).  Included by hala.hpp; the
// no-engine solver forms are in hala_solvers.hpp.
//
// Beyond the reference's type, engined_vector forwards data(), size() and
// resize() to the vector it binds, so it can be handed to any engine call.
#pragma once

namespace hala {

template <class E>
struct is_engine : std::false_type {};
template <>
struct is_engine<gpu_engine> : std::true_type {};
template <>
struct is_engine<mixed_engine> : std::true_type {};

namespace detail {
// the bound vector: a reference for vec = V& / V const&, an owned V otherwise
template <class vec>
struct engined_binding {
    using type = std::remove_reference_t<vec>;
    using variable = std::conditional_t<std::is_reference<vec>::value, std::reference_wrapper<type>, type>;
};
template <class V>
V &get_vec(std::reference_wrapper<V> const &w) {
    return w.get();
}
template <class V>
V &get_vec(V &v) {
    return v;
}
template <class V>
V const &get_vec(V const &v) {
    return v;
}

// vectors the gpu_engine's kernels can read (device memory)
template <class V>
struct is_device_vector : std::false_type {};
template <typename T>
struct is_device_vector<gpu_vector<T>> : std::true_type {};
template <typename T>
struct is_device_vector<gpu_wrapped_array<T>> : std::true_type {};
template <typename T, class VL>
struct is_device_vector<binded_gpu_vector<T, VL>> : std::true_type {};

// a fresh owned vector of the engine's default type (get_vdefault in the reference)
template <class cengine, typename T>
struct vdefault;
template <typename T>
struct vdefault<gpu_engine, T> {
    using type = gpu_vector<T>;
    static type make(gpu_engine const &e) { return type(e.device()); }
};
template <typename T>
struct vdefault<mixed_engine, T> {
    using type = std::vector<T>;
    static type make(mixed_engine const &) { return type(); }
};
}  // namespace detail

template <class cengine, class vec>
struct engined_vector {
    using vtype = typename detail::engined_binding<vec>::type;
    using value_type = scalar_of<vtype>;
    static_assert(is_engine<cengine>::value, "engined_vector: cengine is not an engine");
    static_assert(!std::is_same<cengine, gpu_engine>::value ||
                      detail::is_device_vector<std::remove_cv_t<vtype>>::value,
                  "a vector bound to the gpu_engine must live in device memory (gpu_vector, a wrapped device "
                  "array or gpu_bind_vector); load host data with engine.load() or bind it to a mixed_engine");

    std::reference_wrapper<cengine const> e;
    typename detail::engined_binding<vec>::variable v;

    // non-owning: binds x
    engined_vector(cengine const &engine, vtype &x) : e(engine), v(x) {
        static_assert(std::is_reference<vec>::value, "engined_vector: instantiate with vec& to bind a vector");
    }
    // owning: a fresh empty vector of the engine's default type
    explicit engined_vector(cengine const &engine)
        : e(engine), v(detail::vdefault<cengine, value_type>::make(engine)) {
        static_assert(!std::is_reference<vec>::value, "engined_vector: an owning vector is instantiated without &");
    }

    cengine const &engine() const { return e.get(); }
    vtype const &vector() const { return detail::get_vec(v); }
    vtype &vector() { return detail::get_vec(v); }

    size_t size() const { return get_size(vector()); }
    void resize(size_t new_size) { set_size(new_size, vector()); }
    auto data() { return vector().data(); }
    auto data() const { return vector().data(); }

    // engines are compatible when they have the same type and, on the GPU, the same device
    template <class oengine, class vec2>
    bool engine_match(engined_vector<oengine, vec2> const &) const {
        return false;
    }
    template <class vec2>
    bool engine_match(engined_vector<cengine, vec2> const &other) const {
        return device_of(engine()) == device_of(other.engine());
    }

  private:
    static int device_of(gpu_engine const &g) { return g.device(); }
    static int device_of(mixed_engine const &m) { return m.gpu().device(); }
};

template <class V>
struct is_engined_vector : std::false_type {};
template <class cengine, class vec>
struct is_engined_vector<engined_vector<cengine, vec>> : std::true_type {};
template <class V>
using enable_if_engined = std::enable_if_t<is_engined_vector<std::remove_cv_t<std::remove_reference_t<V>>>::value>;

template <class ev1, class ev2>
bool check_engines(ev1 const &x, ev2 const &y) {
    return x.engine_match(y);
}
template <class ev1, class ev2, class ev3, class... evs>
bool check_engines(ev1 const &x, ev2 const &y, ev3 const &z, evs const &...more) {
    return check_engines(x, y) && check_engines(y, z, more...);
}

// bind_engine_vector(engine, x) / (engined_vector, x) / (non-engine, x) -> x itself
template <class cengine, class vec>
std::enable_if_t<is_engine<cengine>::value, engined_vector<cengine, vec &>> bind_engine_vector(cengine const &e,
                                                                                            vec &x) {
    return engined_vector<cengine, vec &>(e, x);
}
template <class cengine, class vec>
std::enable_if_t<is_engine<cengine>::value, engined_vector<cengine, vec const &>> bind_engine_vector(
    cengine const &e, vec const &x) {
    return engined_vector<cengine, vec const &>(e, x);
}
template <class cengine, class cvec, class vec>
auto bind_engine_vector(engined_vector<cengine, cvec> const &s, vec &x) {
    return bind_engine_vector(s.engine(), x);
}
template <class cvec, class vec>
std::enable_if_t<!is_engine<cvec>::value && !is_engined_vector<cvec>::value, vec &> bind_engine_vector(cvec const &,
                                                                                                       vec &x) {
    return x;
}

template <class cengine, class vec>
cengine const &get_engine(engined_vector<cengine, vec> const &x) {
    return x.engine();
}

namespace detail {
template <class cengine, class vec>
struct unbind<engined_vector<cengine, vec>> {
    static auto &get(engined_vector<cengine, vec> &v) { return v.vector(); }
};
template <class cengine, class vec>
struct unbind<engined_vector<cengine, vec> const> {
    static auto &get(engined_vector<cengine, vec> const &v) { return v.vector(); }
};
}  // namespace detail

// new_vector(x): an owned vector on x's engine, of x's scalar type
template <class cengine, class vec>
auto new_vector(engined_vector<cengine, vec> const &x) {
    using T = typename engined_vector<cengine, vec>::value_type;
    return engined_vector<cengine, typename detail::vdefault<cengine, T>::type>(x.engine());
}
template <class cengine, class vec>
void set_zero(size_t num_entries, engined_vector<cengine, vec> &x) {
    set_zero(x.engine(), num_entries, x.vector());
}

// ---- BLAS-1 (wax/hala_noengine_overloads.hpp) ------------------------------------
template <class VX, class VY, class = enable_if_engined<VX>>
void vcopy(VX const &x, VY &&y) {
    vcopy(x.engine(), detail::raw(x), detail::raw(y));
}
template <class VX, class VY, class = enable_if_engined<VX>>
void vcopy(int N, VX const &x, int incx, VY &&y, int incy) {
    vcopy(x.engine(), N, detail::raw(x), incx, detail::raw(y), incy);
}
template <class VX, class VY, class = enable_if_engined<VX>>
auto dot(VX const &x, VY const &y) {
    return dot(x.engine(), detail::raw(x), detail::raw(y));
}
template <class VX, class VY, class = enable_if_engined<VX>>
auto dotu(VX const &x, VY const &y) {
    return dotu(x.engine(), detail::raw(x), detail::raw(y));
}
template <class VX, class = enable_if_engined<VX>>
auto norm2(VX const &x) {
    return norm2(x.engine(), detail::raw(x));
}
template <typename FP, class VX, class VY, class = enable_if_engined<VX>>
void axpy(FP alpha, VX const &x, VY &&y) {
    axpy(x.engine(), alpha, detail::raw(x), detail::raw(y));
}
template <typename FP, class VX, class = enable_if_engined<VX>>
void scal(FP alpha, VX &&x) {
    scal(x.engine(), alpha, detail::raw(x));
}

// ---- sparse matrices and ILU -----------------------------------------------------
template <class VP, class VI, class VV, class = enable_if_engined<VP>>
auto make_sparse_matrix(int num_rows, int num_cols, int num_nz, VP const &pntr, VI const &indx, VV const &vals) {
    return make_sparse_matrix(pntr.engine(), num_rows, num_cols, num_nz, detail::raw(pntr), detail::raw(indx),
                              detail::raw(vals));
}
template <class VP, class VI, class VV, class = enable_if_engined<VP>>
auto make_sparse_matrix(int num_cols, VP const &pntr, VI const &indx, VV const &vals) {
    return make_sparse_matrix(pntr.engine(), num_cols, detail::raw(pntr), detail::raw(indx), detail::raw(vals));
}
template <typename FPa, class VP, class VI, class VV, class VX, typename FPb, class VY, class = enable_if_engined<VP>>
void sparse_gemv(char trans, int M, int N, FPa alpha, VP const &pntr, VI const &indx, VV const &vals, VX const &x,
                 FPb beta, VY &&y) {
    sparse_gemv(pntr.engine(), trans, M, N, alpha, detail::raw(pntr), detail::raw(indx), detail::raw(vals),
                detail::raw(x), beta, detail::raw(y));
}
template <class VP, class VI, class VV, class = enable_if_engined<VP>>
auto make_triangular_matrix(char uplo, char diag, VP const &pntr, VI const &indx, VV const &vals,
                            char policy = 'N') {
    return make_triangular_matrix(pntr.engine(), uplo, diag, detail::raw(pntr), detail::raw(indx), detail::raw(vals),
                                  policy);
}
template <class VP, class VI, class VV, class = enable_if_engined<VP>>
auto make_ilu(VP const &pntr, VI const &indx, VV const &vals, char policy = 'N') {
    return make_ilu(pntr.engine(), detail::raw(pntr), detail::raw(indx), detail::raw(vals), policy);
}

}  // namespace hala
