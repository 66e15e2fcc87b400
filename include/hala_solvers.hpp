// hala_solvers.hpp -- drop-in hala:: Krylov solvers on the gpu_engine
// (hex/solvers/hala_solvers_core.hpp, hala_solvers_cg.hpp, hala_solvers_gmres.hpp).
//
// solve_cg_ilu / solve_gmres_ilu run entirely on the device (one host
// synchronisation per batch of passes / per restart cycle).  solve_cg and
// solve_gmres with a user preconditioner follow the reference's host-driven
// loop over engine BLAS calls with the same step order.
#pragma once

#include <cmath>
#include <limits>
#include <vector>

#include "hala.hpp"

namespace hala {

// hex/solvers/hala_solvers_core.hpp:52-71
template <typename precision>
struct stop_criteria {
    int max_iter;
    precision tol;
    stop_criteria(int max_iterations) : max_iter(max_iterations), tol(0.0) {}
    template <typename U, typename = std::enable_if_t<std::is_floating_point<U>::value>>
    stop_criteria(U const &tolerance) : max_iter(std::numeric_limits<int>::max()), tol((precision)tolerance) {}
    template <typename U>
    stop_criteria(U const &tolerance, int max_iterations) : max_iter(max_iterations), tol((precision)tolerance) {}
};

// hex/solvers/hala_solvers_core.hpp:80-143: a preconditioner r = P^-1 x over the
// engine's default vectors; the gpu_engine and the mixed_engine both use device
// vectors (the mixed solvers load the data once and work on the GPU).
template <class engine, class VectorLike>
struct define_preconditioner;
template <class VectorLike>
struct define_preconditioner<gpu_engine, VectorLike> {
    using type = std::function<void(gpu_vector<scalar_of<VectorLike>> const &x, gpu_vector<scalar_of<VectorLike>> &r)>;
};
template <class VectorLike>
struct define_preconditioner<mixed_engine, VectorLike> {
    using type = typename define_preconditioner<gpu_engine, VectorLike>::type;
};
template <class engine, class VectorLike>
using preconditioner = typename define_preconditioner<engine, VectorLike>::type;

// No-engine form (hex/solvers/hala_solvers_core.hpp:120-143): the vectors are
// the bound owned vectors new_vector() makes for VectorLike.
template <class VectorLike>
using preconditioner_noe = std::function<void(decltype(new_vector(std::declval<VectorLike>())) const &x,
                                              decltype(new_vector(std::declval<VectorLike>())) &r)>;

// hex/solvers/hala_solvers_core.hpp:159-165 (ROCm: plain int)
template <class compute_engine>
using hala_int = int;

namespace detail {
// temporaries of the host-driven loops: the engine's device vector, or a
// bound owned vector for the no-engine form (its preconditioner receives those)
template <class TV>
struct temp_of;
template <typename T>
struct temp_of<gpu_vector<T>> {
    static gpu_vector<T> make(gpu_engine const &e) { return gpu_vector<T>(e.device()); }
};
template <typename T>
struct temp_of<engined_vector<gpu_engine, gpu_vector<T>>> {
    static engined_vector<gpu_engine, gpu_vector<T>> make(gpu_engine const &e) {
        return engined_vector<gpu_engine, gpu_vector<T>>(e);
    }
};
}  // namespace detail

template <class VX>
void prepare_x(gpu_engine const &engine, int n, VX &x) {
    if ((int)x.size() < n) {
        x.resize(n);
        set_zero(engine, (size_t)n, x);
    }
}

// ---- CG ------------------------------------------------------------------
// solve_cg(engine, stop, pntr, indx, vals, precon, b, x):
// hex/solvers/hala_solvers_cg.hpp:181-244 with solve_cg_core :92-156.
// The host loop runs over temporaries of type TV (see temp_of): device vectors
// for the engine form, bound vectors for the no-engine form.
namespace detail {
template <class TV, class VP, class VI, class VV, class P, class VB, class VX>
int cg_loop(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr,
            VI const &indx, VV const &vals, P &precon, VB const &b, VX &x) {
    using T = scalar_of<VV>;
    const int n = (int)pntr.size() - 1;
    runtime_assert(n > 0 && (int)b.size() == n, "solve_cg: size mismatch");
    prepare_x(engine, n, x);
    auto matrix = make_sparse_matrix(engine, n, pntr, indx, vals);
    TV r = temp_of<TV>::make(engine), p = temp_of<TV>::make(engine), ap = temp_of<TV>::make(engine),
       z = temp_of<TV>::make(engine);
    vcopy(engine, b, r);
    matrix.gemv('N', 1.0, x, 0.0, p);
    axpy(engine, -1.0, p, r);
    precon(r, z);
    vcopy(engine, z, p);
    T zr = dot(engine, r, z);  // conjugated for complex vectors
    int iterations = 1;
    for (;;) {
        matrix.gemv('N', 1.0, p, 0.0, ap);
        iterations++;
        T nzr = dot(engine, p, ap);
        T a = zr / nzr;
        axpy(engine, a, p, x);
        axpy(engine, -a, ap, r);
        if (iterations == stop.max_iter || norm2(engine, r) < stop.tol) break;
        precon(r, z);
        nzr = dot(engine, r, z);
        a = nzr / zr;
        scal(engine, a, p);
        axpy(engine, 1.0, z, p);
        zr = nzr;
    }
    return iterations;
}
}  // namespace detail

template <class VP, class VI, class VV, class VB, class VX>
int solve_cg(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr,
             VI const &indx, VV const &vals, preconditioner<gpu_engine, VX> precon, VB const &b, VX &x) {
    return detail::cg_loop<gpu_vector<scalar_of<VV>>>(engine, stop, pntr, indx, vals, precon, b, x);
}

// solve_cg_ilu(engine, stop, pntr, indx, vals, ilu, b, x): hex/solvers/hala_solvers_cg.hpp:312-326.
// Runs device-resident (hala_cg_f64); identical arithmetic to solve_cg with
// the ILU apply as preconditioner.
// Other scalar types run solve_cg's host loop with the ILU apply as preconditioner.
namespace detail {
template <class VP, class VI, class VV, class VB, class VX>
int cg_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
           VV const &vals, gpu_ilu<double> const &ilu, VB const &b, VX &x, std::true_type) {
    const int n = (int)pntr.size() - 1;
    runtime_assert(n > 0 && (int)b.size() == n, "solve_cg_ilu: size mismatch");
    prepare_x(engine, n, x);
    auto matrix = make_sparse_matrix(engine, n, pntr, indx, vals);
    int it = 0;
    check_hala(hala_cg_f64(matrix.handle(), (double const *)vals.data(), 1, ilu.handle(), nullptr,
                           (double const *)b.data(), (double *)x.data(), stop.tol, stop.max_iter, &it, nullptr, 0),
               "solve_cg_ilu()");
    return it;
}
template <class VP, class VI, class VV, class T, class VB, class VX>
int cg_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
           VV const &vals, gpu_ilu<T> const &ilu, VB const &b, VX &x, std::false_type) {
    return solve_cg(engine, stop, pntr, indx, vals,
                    [&](gpu_vector<T> const &in, gpu_vector<T> &out) { ilu.apply(in, out); }, b, x);
}
}  // namespace detail

template <class VP, class VI, class VV, class VB, class VX>
int solve_cg_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                 VV const &vals, gpu_ilu<scalar_of<VV>> const &ilu, VB const &b, VX &x) {
    return detail::cg_ilu(engine, stop, pntr, indx, vals, ilu, b, x, std::is_same<scalar_of<VV>, double>());
}

// solve_cg_ilu(engine, stop, pntr, indx, vals, b, x, policy): factor on the fly (:354-361).
template <class VP, class VI, class VV, class VB, class VX>
int solve_cg_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                 VV const &vals, VB const &b, VX &x, char policy = 'N') {
    auto ilu = make_ilu(engine, pntr, indx, vals, policy);
    return solve_cg_ilu(engine, stop, pntr, indx, vals, ilu, b, x);
}

// ---- multi-RHS CG with one shared matrix -----------------------------------
// B, X hold nrhs right-hand sides / solutions of n entries each, column-major
// (column k at offset k n), as the reference's sparse_gemm.
template <class V>
gpu_wrapped_array<scalar_of<V>> batch_column(gpu_engine const &engine, V &v, int n, int k) {
    return wrap_gpu_array((scalar_of<V> *)v.data() + (size_t)k * n, (size_t)n, engine.device());
}

// solve_batch_cg(engine, stop, pntr, indx, vals, precon, B, X):
// hex/solvers/hala_solvers_cg_batch.hpp:68-136 (solve_cg_core with batch_dot,
// vdivide, batch_axpy, batch_scal and batch_max_norm2 of
// wax/hala_blas_extensions.hpp).  The preconditioner receives the whole batch.
// Host loop over the engine's per-column BLAS, the reference's step order.
namespace detail {
template <class TV, class VP, class VI, class VV, class P, class VB, class VX>
int batch_cg_loop(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr,
                  VI const &indx, VV const &vals, P &precon, VB const &B, VX &X) {
    using T = scalar_of<VV>;
    const int n = (int)pntr.size() - 1;
    runtime_assert(n > 0 && B.size() % (size_t)n == 0, "solve_batch_cg: size mismatch");
    const int nrhs = (int)(B.size() / (size_t)n);
    prepare_x(engine, (int)B.size(), X);
    auto matrix = make_sparse_matrix(engine, n, pntr, indx, vals);
    TV r = temp_of<TV>::make(engine), p = temp_of<TV>::make(engine), ap = temp_of<TV>::make(engine),
       z = temp_of<TV>::make(engine);
    for (TV *v : {&r, &p, &ap, &z}) v->resize(B.size());
    auto col = [&](auto &v, int k) { return batch_column(engine, v, n, k); };
    vcopy(engine, B, r);
    for (int k = 0; k < nrhs; k++) matrix.gemv('N', 1.0, col(X, k), 0.0, col(p, k));
    axpy(engine, -1.0, p, r);
    precon(r, z);
    vcopy(engine, z, p);
    std::vector<T> zr(nrhs), a(nrhs);
    for (int k = 0; k < nrhs; k++) zr[k] = dot(engine, col(r, k), col(z, k));
    int iterations = 1;
    for (;;) {
        for (int k = 0; k < nrhs; k++) matrix.gemv('N', 1.0, col(p, k), 0.0, col(ap, k));
        iterations++;
        for (int k = 0; k < nrhs; k++) a[k] = zr[k] / dot(engine, col(p, k), col(ap, k));
        for (int k = 0; k < nrhs; k++) axpy(engine, T(1) * a[k], col(p, k), col(X, k));
        for (int k = 0; k < nrhs; k++) axpy(engine, T(-1) * a[k], col(ap, k), col(r, k));
        auto mx = std::real(T(0));  // batch_max_norm2: the largest <r_k, r_k>, then its square root
        for (int k = 0; k < nrhs; k++) {
            const auto d = std::real(dot(engine, col(r, k), col(r, k)));
            if (k == 0 || d > mx) mx = d;
        }
        if (iterations == stop.max_iter || std::sqrt(mx) < stop.tol) break;
        precon(r, z);
        for (int k = 0; k < nrhs; k++) {
            const T nzr = dot(engine, col(r, k), col(z, k));
            scal(engine, nzr / zr[k], col(p, k));
            zr[k] = nzr;
        }
        axpy(engine, 1.0, z, p);
    }
    return iterations;
}
}  // namespace detail

template <class VP, class VI, class VV, class VB, class VX>
int solve_batch_cg(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr,
                   VI const &indx, VV const &vals, preconditioner<gpu_engine, VX> precon, VB const &B, VX &X) {
    return detail::batch_cg_loop<gpu_vector<scalar_of<VV>>>(engine, stop, pntr, indx, vals, precon, B, X);
}

// solve_batch_cg_ilu(engine, stop, pntr, indx, vals, ilu, B, X):
// hex/solvers/hala_solvers_cg_batch.hpp:156-200.  Runs device-resident
// (hala_cg_batch_f64), the same arithmetic as solve_batch_cg with the ILU
// applied to every column.
namespace detail {
template <class VP, class VI, class VV, class VB, class VX>
int batch_cg_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                 VV const &vals, gpu_ilu<double> const &ilu, VB const &B, VX &X, std::true_type) {
    const int n = (int)pntr.size() - 1;
    runtime_assert(n > 0 && B.size() % (size_t)n == 0, "solve_batch_cg_ilu: size mismatch");
    prepare_x(engine, (int)B.size(), X);
    auto matrix = make_sparse_matrix(engine, n, pntr, indx, vals);
    int it = 0;
    check_hala(hala_cg_batch_f64(matrix.handle(), (double const *)vals.data(), 1, ilu.handle(),
                                 (int)(B.size() / (size_t)n), (double const *)B.data(), (double *)X.data(), stop.tol,
                                 stop.max_iter, &it, nullptr, 0),
               "solve_batch_cg_ilu()");
    return it;
}
template <class VP, class VI, class VV, class T, class VB, class VX>
int batch_cg_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                 VV const &vals, gpu_ilu<T> const &ilu, VB const &B, VX &X, std::false_type) {
    const int nrhs = (int)(B.size() / (pntr.size() - 1));
    return solve_batch_cg(engine, stop, pntr, indx, vals,
                          [&](gpu_vector<T> const &in, gpu_vector<T> &out) { ilu.apply(in, out, nrhs); }, B, X);
}
}  // namespace detail

template <class VP, class VI, class VV, class VB, class VX>
int solve_batch_cg_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                       VV const &vals, gpu_ilu<scalar_of<VV>> const &ilu, VB const &B, VX &X) {
    return detail::batch_cg_ilu(engine, stop, pntr, indx, vals, ilu, B, X, std::is_same<scalar_of<VV>, double>());
}

// solve_batch_cg_ilu(engine, stop, pntr, indx, vals, B, X, policy): factor on the fly (:238-249).
template <class VP, class VI, class VV, class VB, class VX>
int solve_batch_cg_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                       VV const &vals, VB const &B, VX &X, char policy = 'N') {
    auto ilu = make_ilu(engine, pntr, indx, vals, policy);
    return solve_batch_cg_ilu(engine, stop, pntr, indx, vals, ilu, B, X);
}

// Device-resident CG with the identity (z = r) or Jacobi (z = r ./ d) preconditioner.
template <class VP, class VI, class VV, class VB, class VX>
int solve_cg_device(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                    VV const &vals, double const *jacobi_diag, VB const &b, VX &x) {
    const int n = (int)pntr.size() - 1;
    prepare_x(engine, n, x);
    auto matrix = make_sparse_matrix(engine, n, pntr, indx, vals);
    int it = 0;
    check_hala(hala_cg_f64(matrix.handle(), (double const *)vals.data(), jacobi_diag ? 2 : 0, nullptr, jacobi_diag,
                           (double const *)b.data(), (double *)x.data(), stop.tol, stop.max_iter, &it, nullptr, 0),
               "solve_cg_device()");
    return it;
}

// ---- GMRES ----------------------------------------------------------------
namespace detail {
// reference BLAS drotg / drot / dtpsv('U','N','N') on the host (the reference
// calls them on the CPU even with a GPU engine, hex/solvers/hala_solvers_gmres.hpp:196-222)
template <typename R, typename = std::enable_if_t<std::is_floating_point<R>::value>>
void rotg(R &a, R &b, R &c, R &s) {
    R sa = a, sb = b;
    R roe = (std::fabs(sa) > std::fabs(sb)) ? sa : sb;
    R scale = std::fabs(sa) + std::fabs(sb), r, z;
    if (scale == R(0)) {
        c = R(1);
        s = R(0);
        r = R(0);
        z = R(0);
    } else {
        R as = sa / scale, bs = sb / scale;
        r = scale * std::sqrt(as * as + bs * bs);
        r = (roe >= R(0)) ? r : -r;
        c = sa / r;
        s = sb / r;
        z = R(1);
        if (std::fabs(sa) > std::fabs(sb)) z = s;
        if (std::fabs(sb) >= std::fabs(sa) && c != R(0)) z = R(1) / c;
    }
    a = r;
    b = z;
}
template <typename R, typename = std::enable_if_t<std::is_floating_point<R>::value>>
void rot(R &x, R &y, R c, R s) {
    R t = c * x + s * y;
    y = c * y - s * x;
    x = t;
}
// reference BLAS crotg / zrotg (real cosine, complex sine) and crot / zrot
template <typename R>
void rotg(std::complex<R> &a, std::complex<R> b, R &c, std::complex<R> &s) {
    const R aa = std::abs(a);
    if (aa == R(0)) {
        c = R(0);
        s = std::complex<R>(1);
        a = b;
        return;
    }
    const R scale = aa + std::abs(b);
    const R norm = scale * std::sqrt(std::norm(a / scale) + std::norm(b / scale));
    const std::complex<R> alpha = a / aa;
    c = aa / norm;
    s = alpha * std::conj(b) / norm;
    a = alpha * norm;
}
template <typename R>
void rot(std::complex<R> &x, std::complex<R> &y, R c, std::complex<R> s) {
    const std::complex<R> t = c * x + s * y;
    y = c * y - std::conj(s) * x;
    x = t;
}
template <typename T>
void tpsv_un(int n, std::vector<T> const &ap, std::vector<T> &x) {
    int kk = n * (n + 1) / 2 - 1;
    for (int j = n - 1; j >= 0; j--) {
        if (x[j] != T(0)) {
            x[j] = x[j] / ap[kk];
            T t = x[j];
            int k = kk - 1;
            for (int i = j - 1; i >= 0; i--) {
                x[i] = x[i] - t * ap[k];
                k--;
            }
        }
        kk -= j + 1;
    }
}
// Krylov projection: coeffs = W_k^H r, r -= W_k coeffs (classical Gram-Schmidt,
// hex/solvers/hala_solvers_gmres.hpp:47-77); FP64 as two dense gemv calls
template <class VR>
std::vector<double> project(gpu_engine const &engine, int n, int k, gpu_vector<double> &W, VR &r,
                            gpu_vector<double> &cd) {
    auto Wk = wrap_gpu_array(W.data(), (size_t)n * k, engine.device());
    cd.resize(k);
    gemv(engine, 'T', n, k, 1.0, Wk, r, 0.0, cd);
    gemv(engine, 'N', n, k, -1.0, Wk, cd, 1.0, r);
    return cd.unload();
}
template <typename T, class VR>
std::vector<T> project(gpu_engine const &engine, int n, int k, gpu_vector<T> &W, VR &r, gpu_vector<T> &) {
    std::vector<T> c(k);
    for (int i = 0; i < k; i++) c[i] = dot(engine, wrap_gpu_array(W.data() + (size_t)i * n, (size_t)n, engine.device()), r);
    for (int i = 0; i < k; i++)
        axpy(engine, -c[i], wrap_gpu_array(W.data() + (size_t)i * n, (size_t)n, engine.device()), r);
    return c;
}
// x += W_k z
template <class VX>
void combine(gpu_engine const &engine, int n, gpu_vector<double> &W, std::vector<double> const &Z,
             gpu_vector<double> &zd, VX &x) {
    zd.load(Z);
    auto Wk = wrap_gpu_array(W.data(), (size_t)n * Z.size(), engine.device());
    gemv(engine, 'N', n, (int)Z.size(), 1.0, Wk, zd, 1.0, x);
}
template <typename T, class VX>
void combine(gpu_engine const &engine, int n, gpu_vector<T> &W, std::vector<T> const &Z, gpu_vector<T> &, VX &x) {
    for (size_t i = 0; i < Z.size(); i++)
        axpy(engine, Z[i], wrap_gpu_array(W.data() + i * n, (size_t)n, engine.device()), x);
}
}  // namespace detail

// solve_gmres(engine, stop, restart, pntr, indx, vals, precon, b, x):
// hex/solvers/hala_solvers_gmres.hpp:125-230 (host Givens, device vectors).
namespace detail {
template <class TV, class VP, class VI, class VV, class P, class VB, class VX>
int gmres_loop(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, int restart,
               VP const &pntr, VI const &indx, VV const &vals, P &precon, VB const &b, VX &x) {
    using T = scalar_of<VV>;
    using R = decltype(std::abs(T(0)));
    const int n = (int)pntr.size() - 1;
    runtime_assert(n > 0 && (int)b.size() == n && restart > 0, "solve_gmres: bad arguments");
    prepare_x(engine, n, x);
    auto matrix = make_sparse_matrix(engine, n, pntr, indx, vals);
    std::vector<T> H, S, Z;
    std::vector<R> C;
    R inner_res = 0, outer_res = R(stop.tol + 1.0);
    int total = 0, outer = 0;
    TV t = temp_of<TV>::make(engine), r = temp_of<TV>::make(engine);
    t.resize(n);
    r.resize(n);
    gpu_vector<T> W((size_t)n * restart, engine.device());
    gpu_vector<T> cd(engine.device()), zd(engine.device());
    while (outer_res > stop.tol && outer < stop.max_iter) {
        H.clear();
        S.clear();
        C.clear();
        Z.clear();
        vcopy(engine, b, t);
        matrix.gemv('N', -1.0, x, 1.0, t);
        precon(t, r);
        total++;
        inner_res = norm2(engine, r);
        scal(engine, R(1) / inner_res, r);
        Z.push_back(T(inner_res));
        detail::copy_(engine.handle(), n, (T const *)r.data(), (T *)W.data());
        int k = 0;
        while (inner_res > stop.tol && k < restart) {
            matrix.gemv('N', 1.0, r, 0.0, t);
            precon(t, r);
            total++;
            std::vector<T> coeffs = detail::project(engine, n, k + 1, W, r, cd);
            R nrm = norm2(engine, r);
            scal(engine, R(1) / nrm, r);
            for (int i = 0; i < k; i++) detail::rot(coeffs[i], coeffs[i + 1], C[i], S[i]);
            T s, beta = T(nrm);
            R c;
            detail::rotg(coeffs[k], beta, c, s);
            H.insert(H.end(), coeffs.begin(), coeffs.end());
            S.push_back(s);
            C.push_back(c);
            inner_res = std::abs(S.back() * Z.back());
            k++;
            if (inner_res > stop.tol && k < restart) {
                detail::copy_(engine.handle(), n, (T const *)r.data(), (T *)W.data() + (size_t)k * n);
                Z.push_back(T(0));
                detail::rot(Z[k - 1], Z[k], C.back(), S.back());
            }
        }
        if (!H.empty()) {
            detail::tpsv_un((int)Z.size(), H, Z);
            detail::combine(engine, n, W, Z, zd, x);
        }
        outer++;
        outer_res = inner_res;
    }
    return total;
}
}  // namespace detail

template <class VP, class VI, class VV, class VB, class VX>
int solve_gmres(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, int restart,
                VP const &pntr, VI const &indx, VV const &vals, preconditioner<gpu_engine, VX> precon, VB const &b,
                VX &x) {
    return detail::gmres_loop<gpu_vector<scalar_of<VV>>>(engine, stop, restart, pntr, indx, vals, precon, b, x);
}

// solve_gmres_ilu(engine, stop, restart, pntr, indx, vals, ilu, b, x):
// hex/solvers/hala_solvers_gmres.hpp:250-267, device-resident (hala_gmres_f64).
namespace detail {
template <class VP, class VI, class VV, class VB, class VX>
int gmres_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr,
              VI const &indx, VV const &vals, gpu_ilu<double> const &ilu, VB const &b, VX &x, std::true_type) {
    const int n = (int)pntr.size() - 1;
    runtime_assert(n > 0 && (int)b.size() == n, "solve_gmres_ilu: size mismatch");
    prepare_x(engine, n, x);
    auto matrix = make_sparse_matrix(engine, n, pntr, indx, vals);
    int total = 0;
    check_hala(hala_gmres_f64(matrix.handle(), (double const *)vals.data(), 1, ilu.handle(),
                              (double const *)b.data(), (double *)x.data(), stop.tol, restart, stop.max_iter, &total,
                              nullptr, 0, nullptr),
               "solve_gmres_ilu()");
    return total;
}
template <class VP, class VI, class VV, class T, class VB, class VX>
int gmres_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr,
              VI const &indx, VV const &vals, gpu_ilu<T> const &ilu, VB const &b, VX &x, std::false_type) {
    return solve_gmres(engine, stop, restart, pntr, indx, vals,
                       [&](gpu_vector<T> const &in, gpu_vector<T> &out) { ilu.apply(in, out); }, b, x);
}
}  // namespace detail

template <class VP, class VI, class VV, class VB, class VX>
int solve_gmres_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr,
                    VI const &indx, VV const &vals, gpu_ilu<scalar_of<VV>> const &ilu, VB const &b, VX &x) {
    return detail::gmres_ilu(engine, stop, restart, pntr, indx, vals, ilu, b, x,
                             std::is_same<scalar_of<VV>, double>());
}

template <class VP, class VI, class VV, class VB, class VX>
int solve_gmres_ilu(gpu_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr,
                    VI const &indx, VV const &vals, VB const &b, VX &x, char policy = 'N') {
    auto ilu = make_ilu(engine, pntr, indx, vals, policy);
    return solve_gmres_ilu(engine, stop, restart, pntr, indx, vals, ilu, b, x);
}

// ---- new entry points (no reference counterpart) ------------------------------
// FP32 ILU(0) + FP32 product inside GMRES(restart), FP64 iterative refinement to
// norm2(b - A x) < rtol norm2(b).  Returns the inner preconditioner applications.
template <class VP, class VI, class VV, class VB, class VX>
int solve_gmres_ir(gpu_engine const &engine, double rtol, int restart, VP const &pntr, VI const &indx,
                   VV const &vals, VB const &b, VX &x, double inner_rtol = 1e-4, int max_refine = 50) {
    const int n = (int)pntr.size() - 1;
    prepare_x(engine, n, x);
    auto matrix = make_sparse_matrix(engine, n, pntr, indx, vals);
    gpu_ilu<double> ilu(engine, pntr, indx, vals, 'N', true);
    gpu_vector<float> v32(vals.size(), engine.device());
    check_hala(hala_convert_f64_f32(engine.handle(), (int64_t)vals.size(), (double const *)vals.data(), v32.data()),
               "solve_gmres_ir()");
    int total = 0;
    check_hala(hala_gmres_ir_f64(matrix.handle(), (double const *)vals.data(), v32.data(), ilu.handle(),
                                 (double const *)b.data(), (double *)x.data(), rtol, inner_rtol, restart, max_refine,
                                 &total, nullptr, 0, nullptr),
               "solve_gmres_ir()");
    return total;
}

// ---- mixed engine: host-resident inputs loaded once, the device solve, x copied back
// (hex/solvers/hala_solvers_cg.hpp:249-264, :329-340; hala_solvers_gmres.hpp and
// hala_solvers_cg_batch.hpp, their mixed_engine overloads).  A user preconditioner
// receives device vectors, as with the gpu_engine.
namespace detail {
template <class VP, class VI, class VV, class F>
int mixed_solve(mixed_engine const &engine, VP const &pntr, VI const &indx, VV const &vals, F &&run) {
    auto gp = engine.gpu().load(pntr);
    auto gi = engine.gpu().load(indx);
    auto gv = engine.gpu().load(vals);
    return run(gp, gi, gv);
}
}  // namespace detail

template <class VP, class VI, class VV, class VB, class VX>
int solve_cg(mixed_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
             VV const &vals, preconditioner<mixed_engine, VX> precon, VB const &b, VX &x) {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(b);
        auto gx = gpu_bind_vector(engine.gpu(), x);
        return solve_cg(engine.gpu(), stop, gp, gi, gv, precon, gb, gx);
    });
}
template <class VP, class VI, class VV, class VB, class VX>
int solve_cg_ilu(mixed_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                 VV const &vals, mixed_ilu<scalar_of<VV>> const &ilu, VB const &b, VX &x) {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(b);
        auto gx = gpu_bind_vector(engine.gpu(), x);
        return solve_cg_ilu(engine.gpu(), stop, gp, gi, gv, ilu.cilu(), gb, gx);
    });
}
template <class VP, class VI, class VV, class VB, class VX>
int solve_cg_ilu(mixed_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                 VV const &vals, VB const &b, VX &x, char policy = 'N') {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(b);
        auto gx = gpu_bind_vector(engine.gpu(), x);
        return solve_cg_ilu(engine.gpu(), stop, gp, gi, gv, gb, gx, policy);
    });
}
template <class VP, class VI, class VV, class VB, class VX>
int solve_batch_cg(mixed_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                   VV const &vals, preconditioner<mixed_engine, VX> precon, VB const &B, VX &X) {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(B);
        auto gx = gpu_bind_vector(engine.gpu(), X);
        return solve_batch_cg(engine.gpu(), stop, gp, gi, gv, precon, gb, gx);
    });
}
template <class VP, class VI, class VV, class VB, class VX>
int solve_batch_cg_ilu(mixed_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr,
                       VI const &indx, VV const &vals, mixed_ilu<scalar_of<VV>> const &ilu, VB const &B, VX &X) {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(B);
        auto gx = gpu_bind_vector(engine.gpu(), X);
        return solve_batch_cg_ilu(engine.gpu(), stop, gp, gi, gv, ilu.cilu(), gb, gx);
    });
}
template <class VP, class VI, class VV, class VB, class VX>
int solve_batch_cg_ilu(mixed_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr,
                       VI const &indx, VV const &vals, VB const &B, VX &X, char policy = 'N') {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(B);
        auto gx = gpu_bind_vector(engine.gpu(), X);
        return solve_batch_cg_ilu(engine.gpu(), stop, gp, gi, gv, gb, gx, policy);
    });
}
template <class VP, class VI, class VV, class VB, class VX>
int solve_gmres(mixed_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr,
                VI const &indx, VV const &vals, preconditioner<mixed_engine, VX> precon, VB const &b, VX &x) {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(b);
        auto gx = gpu_bind_vector(engine.gpu(), x);
        return solve_gmres(engine.gpu(), stop, restart, gp, gi, gv, precon, gb, gx);
    });
}
template <class VP, class VI, class VV, class VB, class VX>
int solve_gmres_ilu(mixed_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr,
                    VI const &indx, VV const &vals, mixed_ilu<scalar_of<VV>> const &ilu, VB const &b, VX &x) {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(b);
        auto gx = gpu_bind_vector(engine.gpu(), x);
        return solve_gmres_ilu(engine.gpu(), stop, restart, gp, gi, gv, ilu.cilu(), gb, gx);
    });
}
template <class VP, class VI, class VV, class VB, class VX>
int solve_gmres_ilu(mixed_engine const &engine, stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr,
                    VI const &indx, VV const &vals, VB const &b, VX &x, char policy = 'N') {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(b);
        auto gx = gpu_bind_vector(engine.gpu(), x);
        return solve_gmres_ilu(engine.gpu(), stop, restart, gp, gi, gv, gb, gx, policy);
    });
}
template <class VP, class VI, class VV, class VB, class VX>
int solve_gmres_ir(mixed_engine const &engine, double rtol, int restart, VP const &pntr, VI const &indx,
                   VV const &vals, VB const &b, VX &x, double inner_rtol = 1e-4, int max_refine = 50) {
    return detail::mixed_solve(engine, pntr, indx, vals, [&](auto &gp, auto &gi, auto &gv) {
        auto gb = engine.gpu().load(b);
        auto gx = gpu_bind_vector(engine.gpu(), x);
        return solve_gmres_ir(engine.gpu(), rtol, restart, gp, gi, gv, gb, gx, inner_rtol, max_refine);
    });
}

// ---- no-engine forms: the engine comes from the bound vectors ------------------
// hex/solvers/hala_solvers_cg.hpp:181-227 (solve_cg), :295-305 and :345-352
// (solve_cg_ilu), hala_solvers_cg_batch.hpp and hala_solvers_gmres.hpp (their
// no-engine overloads).  Vectors are bound with bind_engine_vector(); the
// user's preconditioner receives bound vectors (preconditioner_noe).  The
// ILU forms run the engine's device-resident solvers.
namespace detail {
template <class VV>
using engine_of = std::decay_t<decltype(std::declval<VV const &>().engine())>;
template <class VV>
constexpr bool on_gpu_engine() {
    return std::is_same<engine_of<VV>, gpu_engine>::value;
}
}  // namespace detail

template <class VP, class VI, class VV, class VB, class VX, class = enable_if_engined<VV>>
int solve_cg(stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx, VV const &vals,
             preconditioner_noe<VV> precon, VB const &b, VX &x) {
    static_assert(detail::on_gpu_engine<VV>(), "no-engine solve_cg with a preconditioner: bind to the gpu_engine");
    using TV = engined_vector<gpu_engine, gpu_vector<scalar_of<VV>>>;
    return detail::cg_loop<TV>(vals.engine(), stop, detail::raw(pntr), detail::raw(indx), detail::raw(vals), precon,
                               detail::raw(b), detail::raw(x));
}
template <class VP, class VI, class VV, class ILU, class VB, class VX, class = enable_if_engined<VV>>
int solve_cg_ilu(stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx, VV const &vals,
                 ILU const &ilu, VB const &b, VX &x) {
    return solve_cg_ilu(vals.engine(), stop, detail::raw(pntr), detail::raw(indx), detail::raw(vals), ilu,
                        detail::raw(b), detail::raw(x));
}
template <class VP, class VI, class VV, class VB, class VX, class = enable_if_engined<VV>>
int solve_cg_ilu(stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx, VV const &vals,
                 VB const &b, VX &x, char policy = 'N') {
    return solve_cg_ilu(vals.engine(), stop, detail::raw(pntr), detail::raw(indx), detail::raw(vals),
                        detail::raw(b), detail::raw(x), policy);
}
template <class VP, class VI, class VV, class VB, class VX, class = enable_if_engined<VV>>
int solve_batch_cg(stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                   VV const &vals, preconditioner_noe<VV> precon, VB const &B, VX &X) {
    static_assert(detail::on_gpu_engine<VV>(),
                  "no-engine solve_batch_cg with a preconditioner: bind to the gpu_engine");
    using TV = engined_vector<gpu_engine, gpu_vector<scalar_of<VV>>>;
    return detail::batch_cg_loop<TV>(vals.engine(), stop, detail::raw(pntr), detail::raw(indx), detail::raw(vals),
                                     precon, detail::raw(B), detail::raw(X));
}
template <class VP, class VI, class VV, class ILU, class VB, class VX, class = enable_if_engined<VV>>
int solve_batch_cg_ilu(stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                       VV const &vals, ILU const &ilu, VB const &B, VX &X) {
    return solve_batch_cg_ilu(vals.engine(), stop, detail::raw(pntr), detail::raw(indx), detail::raw(vals), ilu,
                              detail::raw(B), detail::raw(X));
}
template <class VP, class VI, class VV, class VB, class VX, class = enable_if_engined<VV>>
int solve_batch_cg_ilu(stop_criteria<get_precision_type<VV>> const &stop, VP const &pntr, VI const &indx,
                       VV const &vals, VB const &B, VX &X, char policy = 'N') {
    return solve_batch_cg_ilu(vals.engine(), stop, detail::raw(pntr), detail::raw(indx), detail::raw(vals),
                              detail::raw(B), detail::raw(X), policy);
}
template <class VP, class VI, class VV, class VB, class VX, class = enable_if_engined<VV>>
int solve_gmres(stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr, VI const &indx,
                VV const &vals, preconditioner_noe<VV> precon, VB const &b, VX &x) {
    static_assert(detail::on_gpu_engine<VV>(), "no-engine solve_gmres with a preconditioner: bind to the gpu_engine");
    using TV = engined_vector<gpu_engine, gpu_vector<scalar_of<VV>>>;
    return detail::gmres_loop<TV>(vals.engine(), stop, restart, detail::raw(pntr), detail::raw(indx),
                                  detail::raw(vals), precon, detail::raw(b), detail::raw(x));
}
template <class VP, class VI, class VV, class ILU, class VB, class VX, class = enable_if_engined<VV>>
int solve_gmres_ilu(stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr, VI const &indx,
                    VV const &vals, ILU const &ilu, VB const &b, VX &x) {
    return solve_gmres_ilu(vals.engine(), stop, restart, detail::raw(pntr), detail::raw(indx), detail::raw(vals), ilu,
                           detail::raw(b), detail::raw(x));
}
template <class VP, class VI, class VV, class VB, class VX, class = enable_if_engined<VV>>
int solve_gmres_ilu(stop_criteria<get_precision_type<VV>> const &stop, int restart, VP const &pntr, VI const &indx,
                    VV const &vals, VB const &b, VX &x, char policy = 'N') {
    return solve_gmres_ilu(vals.engine(), stop, restart, detail::raw(pntr), detail::raw(indx), detail::raw(vals),
                           detail::raw(b), detail::raw(x), policy);
}

}  // namespace hala
