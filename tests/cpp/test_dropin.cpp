// test_dropin.cpp -- user code written against HALA's API, compiled with g++
// against include/ (hala.hpp, hala_solvers.hpp) and libhala_amd.so (no HIP headers).  The cases are
// the reference's own sparse/solver tests (tests/sparse_tests.hpp,
// tests/solvers_tests.hpp, tests/cuda_sparse_tests.hpp) and the GPU example's
// PCG table (examples/example_gpu.cpp:90-98, :363-365).
#include <cmath>
#include <complex>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <valarray>
#include <vector>

#include "hala.hpp"
#include "hala_solvers.hpp"

static int failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);          \
            failures++;                                                          \
        }                                                                        \
    } while (0)

static double diffsum(std::vector<double> const &a, std::vector<double> const &b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); i++) s += std::fabs(a[i] - b[i]);
    return s;
}

static std::vector<double> densemul(std::vector<int> const &p, std::vector<int> const &ix, std::vector<double> const &v,
                                    std::vector<double> const &x, bool trans = false) {
    int n = (int)p.size() - 1;
    std::vector<double> y(trans ? x.size() : n, 0.0);
    for (int r = 0; r < n; r++)
        for (int j = p[r]; j < p[r + 1]; j++) {
            if (trans) y[ix[j]] += v[j] * x[r];
            else y[r] += v[j] * x[ix[j]];
        }
    return y;
}

// tests/solvers_tests.hpp:12-137 for one scalar type: CG, batch CG and GMRES
// with ILU(0) on the reference's 5x5 systems (complex: complex-valued x).
template <typename T>
static double errnorm(std::vector<T> const &x, std::vector<T> const &xref) {
    double e = 0, r = 0;
    for (size_t i = 0; i < xref.size(); i++) {
        e += std::norm(std::complex<double>(x[i]) - std::complex<double>(xref[i]));
        r += std::norm(std::complex<double>(xref[i]));
    }
    return x.size() == xref.size() ? std::sqrt(e / r) : 1e300;
}
template <typename T>
static std::vector<T> tmul(std::vector<int> const &p, std::vector<int> const &ix, std::vector<T> const &v,
                           std::vector<T> const &x, int nrhs = 1) {
    int n = (int)p.size() - 1;
    std::vector<T> y((size_t)n * nrhs, T(0));
    for (int k = 0; k < nrhs; k++)
        for (int r = 0; r < n; r++)
            for (int j = p[r]; j < p[r + 1]; j++) y[k * n + r] += v[j] * x[k * n + ix[j]];
    return y;
}
template <typename T>
static T conjv(T v) {
    return v;
}
template <typename R>
static std::complex<R> conjv(std::complex<R> v) {
    return std::conj(v);
}
template <typename T>
struct mkv {
    static T get(double re, double) { return T(re); }
};
template <typename R>
struct mkv<std::complex<R>> {
    static std::complex<R> get(double re, double im) { return {R(re), R(im)}; }
};
// A user's PCG written once in the no-engine style (no engine argument in any
// call; the engine travels with the bound vectors), as the reference documents
// it (examples/example_gpu.cpp:208-285): matrix and ILU from the bound arrays,
// work vectors from new_vector(), the engine-less BLAS and sparse calls.
template <class VP, class VI, class VV, class VB, class VX>
static int user_pcg(int max_iter, hala::get_precision_type<VV> tol, VP const &pntr, VI const &indx, VV const &vals,
                    VB const &b, VX &x) {
    const int rows = hala::get_size_int(pntr) - 1;
    auto A = hala::make_sparse_matrix(rows, pntr, indx, vals);
    auto M = hala::make_ilu(pntr, indx, vals, 'N');
    auto work = hala::new_vector(x);
    hala::set_size(std::max(hala::ilu_buffer_size(M, b, x, 1), hala::sparse_gemv_buffer_size(A, 'N', 1.0, b, 0.0, x)),
                   work);
    auto res = hala::new_vector(x);
    auto dir = hala::new_vector(x);
    auto adir = hala::new_vector(x);
    auto pres = hala::new_vector(x);
    hala::vcopy(b, res);
    hala::sparse_gemv(A, 'N', -1.0, x, 1.0, res, work);
    hala::ilu_apply(M, res, pres);
    hala::vcopy(pres, dir);
    auto rz = hala::dot(res, pres);
    int it = 1;
    for (;;) {
        hala::sparse_gemv(A, 'N', 1.0, dir, 0.0, adir, work);
        it++;
        auto step = rz / hala::dot(adir, dir);
        hala::axpy(step, dir, x);
        hala::axpy(-step, adir, res);
        if (it >= max_iter || hala::norm2(res) < tol) break;
        hala::ilu_apply(M, res, pres, 1, work);
        auto rz_new = hala::dot(res, pres);
        hala::scal(rz_new / rz, dir);
        hala::axpy(1.0, pres, dir);
        rz = rz_new;
    }
    return it;
}
// ... and its engine overload: bind every array to the engine, run the no-engine body
template <class E, class VP, class VI, class VV, class VB, class VX>
static int user_pcg(E const &engine, int max_iter, hala::get_precision_type<VV> tol, VP const &pntr, VI const &indx,
                    VV const &vals, VB const &b, VX &x) {
    auto bp = hala::bind_engine_vector(engine, pntr);
    auto bi = hala::bind_engine_vector(engine, indx);
    auto bv = hala::bind_engine_vector(engine, vals);
    auto bb = hala::bind_engine_vector(engine, b);
    auto bx = hala::bind_engine_vector(engine, x);
    return user_pcg(max_iter, tol, bp, bi, bv, bb, bx);
}

template <typename T>
static void typed_solvers(hala::gpu_engine const &engine, const char *name, double tol) {
    const bool single = std::is_same<T, float>::value || std::is_same<T, std::complex<float>>::value;
    const double stop = single ? 1.E-4 : 1.E-9;
    std::vector<int> pntr = {0, 2, 5, 8, 11, 13};
    std::vector<int> indx = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4};
    std::vector<T> vals = {2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0};
    std::vector<T> xref(15);
    for (int i = 0; i < 15; i++) xref[i] = mkv<T>::get(1.0 + i, 0.5 * i - 3.0);
    std::vector<T> x5(xref.begin(), xref.begin() + 5);
    auto run = [&](std::vector<int> const &p, std::vector<int> const &ix, std::vector<T> const &v, bool exact) {
        auto gp = hala::make_gpu_vector(p);
        auto gi = hala::make_gpu_vector(ix);
        auto gv = hala::make_gpu_vector(v);
        auto ilu = hala::make_ilu(engine, gp, gi, gv, 'N');
        auto gb = hala::make_gpu_vector(tmul(p, ix, v, x5));
        std::vector<T> x;
        hala::gpu_vector<T> gx;
        // CG with the ILU as preconditioner lambda
        if (exact) {
            hala::solve_cg(engine, 2, gp, gi, gv,
                           [&](hala::gpu_vector<T> const &in, hala::gpu_vector<T> &out) { hala::ilu_apply(ilu, in, out); },
                           gb, gx);
            gx.unload(x);
            CHECK(errnorm(x, x5) < tol);
        }
        // CG and GMRES with the ILU object, then three right-hand sides
        gx = hala::gpu_vector<T>();
        hala::solve_cg_ilu(engine, {stop, 50}, gp, gi, gv, ilu, gb, gx);
        gx.unload(x);
        CHECK(errnorm(x, x5) < tol * 100);
        gx = hala::gpu_vector<T>();
        // the stop criteria are typed by the precision (float for float and complex<float>)
        using P = hala::get_precision_type<std::vector<T>>;
        hala::solve_gmres_ilu(engine, exact ? hala::stop_criteria<P>(2) : hala::stop_criteria<P>(stop), exact ? 2 : 6,
                              gp, gi, gv, ilu, gb, gx);
        gx.unload(x);
        CHECK(errnorm(x, x5) < tol * 100);
        auto gB = hala::make_gpu_vector(tmul(p, ix, v, xref, 3));
        hala::gpu_vector<T> gX;
        hala::solve_batch_cg_ilu(engine, {stop, 50}, gp, gi, gv, ilu, gB, gX);
        gX.unload(x);
        CHECK(errnorm(x, xref) < tol * 100);
    };
    run(pntr, indx, vals, true);  // tri-diagonal: ILU(0) is exact
    pntr = {0, 3, 6, 9, 12, 15};
    indx = {0, 1, 4, 0, 1, 2, 1, 2, 3, 2, 3, 4, 0, 3, 4};
    vals = {2.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 2.0};
    run(pntr, indx, vals, false);
    std::printf("PASS solvers<%s>\n", name);
}

// tests/sparse_tests.hpp:314-366 (sparse_trsm) for one scalar type: B = 2 op(A) Xref,
// then sparse_trsm(transa, transb, nrhs, tri, 0.5, B) must give Xref back
template <typename T>
static void typed_trsm(hala::gpu_engine const &engine, const char *name, double tol) {
    const int M = 3, nrhs = 6;
    const bool cplx = !std::is_same<T, float>::value && !std::is_same<T, double>::value;
    std::vector<T> Adense = {mkv<T>::get(1, 0), mkv<T>::get(4, 0), mkv<T>::get(5, 0), T(0), mkv<T>::get(2, 0), T(0),
                             T(0), T(0), mkv<T>::get(3, 0)};  // column-major lower triangle
    std::vector<T> Xref(M * nrhs);
    for (int i = 0; i < M * nrhs; i++) Xref[i] = mkv<T>::get(-3.0 + i, cplx ? 0.5 * i : 0.0);
    auto opA = [&](int r, int c, char t) -> T {
        T v = (t == 'N') ? Adense[r + c * M] : Adense[c + r * M];
        return t == 'C' ? conjv(v) : v;
    };
    auto run = [&](char uplo, std::vector<int> const &p, std::vector<int> const &ix, std::vector<T> const &v,
                   char transa, char transb, char opchar) {
        // B = 2 op(A_uplo) Xref in the layout transb selects
        std::vector<T> B(M * nrhs);
        for (int k = 0; k < nrhs; k++)
            for (int r = 0; r < M; r++) {
                T s = T(0);
                for (int c = 0; c < M; c++) {
                    const T xc = (transb == 'N') ? Xref[c + k * M] : Xref[k + c * nrhs];
                    s += T(2) * opA(r, c, opchar) * xc;
                }
                if (transb == 'N') B[r + k * M] = s;
                else B[k + r * nrhs] = s;
            }
        auto gp = hala::make_gpu_vector(p);
        auto gi = hala::make_gpu_vector(ix);
        auto gv = hala::make_gpu_vector(v);
        auto tri = hala::make_triangular_matrix(engine, uplo, 'N', gp, gi, gv);
        auto gB = hala::make_gpu_vector(B);
        hala::sparse_trsm(transa, transb, nrhs, tri, T(0.5), gB);
        std::vector<T> X;
        gB.unload(X);
        double err = 0.0;
        for (int i = 0; i < M * nrhs; i++) err += std::abs(std::complex<double>(X[i]) - std::complex<double>(Xref[i]));
        CHECK(err < tol);
    };
    std::vector<int> lp = {0, 1, 3, 5}, li = {0, 0, 1, 0, 2};
    std::vector<T> lv = {mkv<T>::get(1, 0), mkv<T>::get(4, 0), mkv<T>::get(2, 0), mkv<T>::get(5, 0), mkv<T>::get(3, 0)};
    run('L', lp, li, lv, 'N', 'N', 'N');
    run('L', lp, li, lv, 'T', 'N', 'T');
    run('L', lp, li, lv, 'N', 'T', 'N');
    run('L', lp, li, lv, 'C', 'N', 'C');
    run('L', lp, li, lv, 'N', 'C', 'N');
    // the upper triangle A^T: 'N' on it is A^T, 'T' on it is A
    std::vector<int> up = {0, 3, 4, 5}, ui = {0, 1, 2, 1, 2};
    std::vector<T> uv = {mkv<T>::get(1, 0), mkv<T>::get(4, 0), mkv<T>::get(5, 0), mkv<T>::get(2, 0), mkv<T>::get(3, 0)};
    run('U', up, ui, uv, 'N', 'N', 'T');
    run('U', up, ui, uv, 'T', 'N', 'N');
    std::printf("PASS sparse_trsm<%s>\n", name);
    // sparse_gemm on the same lower triangle as a general matrix: C = 2 A Xref,
    // then 'T' on it: C = 2 A^T Xref (tests/sparse_tests.hpp gemm cases)
    for (char ta : {'N', 'T'}) {
        auto gp = hala::make_gpu_vector(lp);
        auto gi = hala::make_gpu_vector(li);
        auto gv = hala::make_gpu_vector(lv);
        auto A = hala::make_sparse_matrix(engine, M, gp, gi, gv);
        auto gX = hala::make_gpu_vector(Xref);
        hala::gpu_vector<T> gC;
        hala::sparse_gemm(A, ta, 'N', M, nrhs, T(2.0), gX, M, T(0.0), gC, M);
        std::vector<T> Cv;
        gC.unload(Cv);
        double err = 0.0;
        for (int k = 0; k < nrhs; k++)
            for (int r = 0; r < M; r++) {
                T s = T(0);
                for (int c = 0; c < M; c++) s += T(2) * opA(r, c, ta) * Xref[c + k * M];
                err += std::abs(std::complex<double>(Cv[r + k * M]) - std::complex<double>(s));
            }
        CHECK(err < tol);
    }
    std::printf("PASS sparse_gemm<%s>\n", name);
}

int main() {
    hala::gpu_engine engine(0);

    // tests/cuda_sparse_tests.hpp:10-30 and tests/sparse_tests.hpp:165-191
    {
        std::vector<int> pntr = {0, 2, 5, 8, 11, 13};
        std::vector<int> indx = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4};
        std::vector<double> vals = {2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0};
        std::vector<double> x = {1.0, 2.0, 3.0, 4.0, 5.0}, y;
        auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
        auto gx = hala::make_gpu_vector(x);
        hala::gpu_vector<double> gy;
        hala::sparse_gemv(engine, 'N', 5, 5, 2.0, gp, gi, gv, gx, 0.0, gy);
        gy.unload(y);
        auto ref = densemul(pntr, indx, vals, x);
        for (auto &r : ref) r *= 2.0;
        CHECK(diffsum(y, ref) == 0.0);
        hala::sparse_gemv(engine, 'T', 5, 5, 2.0, gp, gi, gv, gx, 0.0, gy);
        gy.unload(y);
        ref = densemul(pntr, indx, vals, x, true);
        for (auto &r : ref) r *= 2.0;
        CHECK(diffsum(y, ref) == 0.0);
        auto mat = hala::make_sparse_matrix(engine, 5, gp, gi, gv);
        mat.gemv('N', 1.0, gx, 0.0, gy);
        CHECK(std::fabs(hala::dot(engine, gy, gx) - 190.0) < 1e-12);  // x^T A x = 4+16+36+64+70
        std::printf("PASS sparse_gemv\n");
    }
    // tests/sparse_tests.hpp:259-311 (lower, non-unit; N and T)
    {
        std::vector<int> pntr = {0, 1, 3, 5}, indx = {0, 0, 1, 0, 2};
        std::vector<double> vals = {1.0, 4.0, 2.0, 5.0, 3.0}, xref = {-3.0, -2.0, -1.0};
        // A = [[1,0,0],[4,2,0],[5,0,3]];  b = 2 A xref (N), 2 A^T xref (T); alpha = 0.5
        std::vector<double> bN = {-6.0, -32.0, -36.0}, bT = {-32.0, -8.0, -6.0}, x;
        auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
        auto tri = hala::make_triangular_matrix(engine, 'L', 'N', gp, gi, gv);
        hala::gpu_vector<double> gx;
        hala::sparse_trsv('N', tri, 0.5, hala::make_gpu_vector(bN), gx);
        gx.unload(x);
        CHECK(diffsum(x, xref) < 1e-14);
        hala::sparse_trsv('T', tri, 0.5, hala::make_gpu_vector(bT), gx);
        gx.unload(x);
        CHECK(diffsum(x, xref) < 1e-14);
        std::printf("PASS sparse_trsv\n");
    }
    // tests/sparse_tests.hpp:423-440 (ILU known answer)
    {
        std::vector<int> pntr = {0, 3, 6, 9}, indx = {0, 1, 2, 0, 1, 2, 0, 1, 2};
        std::vector<double> vals = {3.0, 2.0, 1.0, 2.0, 4.0, 3.0, 2.0, 1.0, 6.0}, x = {1.0, 2.0, 3.0}, r;
        auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
        auto ilu = hala::make_ilu(engine, gp, gi, gv, 'N');
        hala::gpu_vector<double> gr;
        hala::ilu_apply(ilu, hala::make_gpu_vector(x), gr);
        gr.unload(r);
        CHECK(diffsum(r, {1.0 / 9.0, 1.0 / 9.0, 4.0 / 9.0}) < 1e-14);
        std::printf("PASS ilu\n");
    }
    // tests/solvers_tests.hpp:12-55 (CG)
    {
        std::vector<int> pntr = {0, 2, 5, 8, 11, 13};
        std::vector<int> indx = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4};
        std::vector<double> vals = {2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0};
        std::vector<double> xref = {1.0, 2.0, 3.0, 4.0, 5.0}, x;
        auto b = densemul(pntr, indx, vals, xref);
        auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
        auto gb = hala::make_gpu_vector(b);
        auto ilu = hala::make_ilu(engine, gp, gi, gv, 'N');
        hala::gpu_vector<double> gx;
        hala::solve_cg(engine, 2, gp, gi, gv,
                       [&](hala::gpu_vector<double> const &inx, hala::gpu_vector<double> &outr) {
                           hala::ilu_apply(ilu, inx, outr);
                       },
                       gb, gx);
        gx.unload(x);
        CHECK(diffsum(x, xref) < 1e-14);
        pntr = {0, 3, 6, 9, 12, 15};
        indx = {0, 1, 4, 0, 1, 2, 1, 2, 3, 2, 3, 4, 0, 3, 4};
        vals = {2.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 2.0};
        b = densemul(pntr, indx, vals, xref);
        gp = hala::make_gpu_vector(pntr);
        gi = hala::make_gpu_vector(indx);
        gv = hala::make_gpu_vector(vals);
        gb = hala::make_gpu_vector(b);
        auto ilu2 = hala::make_ilu(engine, gp, gi, gv, 'N');
        gx.load(std::vector<double>{1.0, 3.0, -4.0, -7.0, 11.0});
        hala::solve_cg_ilu(engine, 1.E-9, gp, gi, gv, ilu2, gb, gx);
        gx.unload(x);
        CHECK(diffsum(x, xref) < 1e-9 * 7.5 * 10);
        gx.load(std::vector<double>{1.0, 3.0, -4.0, -7.0, 11.0});
        hala::solve_cg_ilu(engine, 1.E-9, gp, gi, gv, gb, gx);  // factorisation behind the scene
        std::vector<double> x2;
        gx.unload(x2);
        CHECK(diffsum(x2, x) == 0.0);
        std::printf("PASS solve_cg\n");
    }
    // tests/solvers_tests.hpp:56-104 (batch CG, three right-hand sides)
    {
        const int N = 5, nrhs = 3;
        std::vector<int> pntr = {0, 2, 5, 8, 11, 13};
        std::vector<int> indx = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4};
        std::vector<double> vals = {2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0};
        std::vector<double> xref(N * nrhs), b, x;
        for (int i = 0; i < N * nrhs; i++) xref[i] = 1.0 + 0.5 * i - 0.125 * (i % 4);
        auto batch_rhs = [&]() {
            b.clear();
            for (int k = 0; k < nrhs; k++) {
                auto bk = densemul(pntr, indx, vals, std::vector<double>(xref.begin() + k * N, xref.begin() + (k + 1) * N));
                b.insert(b.end(), bk.begin(), bk.end());
            }
        };
        auto relerr = [&](std::vector<double> const &v) {
            double e = 0, nr = 0;
            for (int i = 0; i < N * nrhs; i++) {
                e += (v[i] - xref[i]) * (v[i] - xref[i]);
                nr += xref[i] * xref[i];
            }
            return std::sqrt(e / nr);
        };
        batch_rhs();
        auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
        auto gb = hala::make_gpu_vector(b);
        auto ilu = hala::make_ilu(engine, gp, gi, gv, 'N');
        hala::gpu_vector<double> gx;
        // tri-diagonal: the ILU is the exact factorisation, convergence is instant
        hala::solve_batch_cg(engine, 2, gp, gi, gv,
                             [&](hala::gpu_vector<double> const &inx, hala::gpu_vector<double> &outr) {
                                 hala::ilu_apply(ilu, inx, outr, nrhs);
                             },
                             gb, gx);
        gx.unload(x);
        CHECK(relerr(x) < 1e-14);
        pntr = {0, 3, 6, 9, 12, 15};
        indx = {0, 1, 4, 0, 1, 2, 1, 2, 3, 2, 3, 4, 0, 3, 4};
        vals = {2.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 2.0};
        batch_rhs();
        gp = hala::make_gpu_vector(pntr);
        gi = hala::make_gpu_vector(indx);
        gv = hala::make_gpu_vector(vals);
        gb = hala::make_gpu_vector(b);
        auto ilu2 = hala::make_ilu(engine, gp, gi, gv, 'N');
        gx = hala::gpu_vector<double>(N * nrhs, 0);
        gx.fill(0.0);
        int it = hala::solve_batch_cg_ilu(engine, {1.E-9, 6}, gp, gi, gv, ilu2, gb, gx);
        gx.unload(x);
        CHECK(it <= 6 && relerr(x) < 1e-8);
        // the host loop with the same ILU applied column by column takes the same steps
        hala::gpu_vector<double> gx2(N * nrhs, 0);
        gx2.fill(0.0);
        int it2 = hala::solve_batch_cg(engine, {1.E-9, 6}, gp, gi, gv,
                                       [&](hala::gpu_vector<double> const &inx, hala::gpu_vector<double> &outr) {
                                           hala::ilu_apply(ilu2, inx, outr, nrhs);
                                       },
                                       gb, gx2);
        std::vector<double> x2;
        gx2.unload(x2);
        CHECK(it2 == it && diffsum(x2, x) == 0.0);
        hala::gpu_vector<double> gx3;
        hala::solve_batch_cg_ilu(engine, 1.E-9, gp, gi, gv, gb, gx3);  // factorisation behind the scene
        std::vector<double> x3;
        gx3.unload(x3);
        CHECK(relerr(x3) < 1e-8);
        std::printf("PASS solve_batch_cg\n");
    }
    // the same cases through the mixed engine: std::vector in and out
    // (tests/sparse_tests.cpp:52-63, tests/solvers_tests.cpp:59-67 run them so)
    {
        hala::mixed_engine mengine(engine);
        std::vector<int> pntr = {0, 2, 5, 8, 11, 13};
        std::vector<int> indx = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4};
        std::vector<double> vals = {2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 1.0, 2.0};
        std::vector<double> xref = {1.0, 2.0, 3.0, 4.0, 5.0}, y, x;
        // BLAS-1 and the product
        std::vector<double> a = {1.0, -2.0, 3.0}, b2 = {4.0, 5.0, -6.0};
        CHECK(hala::dot(mengine, a, b2) == 4.0 - 10.0 - 18.0);
        CHECK(std::fabs(hala::norm2(mengine, a) - std::sqrt(14.0)) < 1e-15);
        hala::axpy(mengine, 2.0, a, b2);
        CHECK(diffsum(b2, {6.0, 1.0, 0.0}) == 0.0);
        hala::scal(mengine, -1.0, b2);
        CHECK(diffsum(b2, {-6.0, -1.0, 0.0}) == 0.0);
        std::vector<double> c2;
        hala::vcopy(mengine, a, c2);
        CHECK(diffsum(c2, a) == 0.0);
        hala::sparse_gemv(mengine, 'N', 5, 5, 2.0, pntr, indx, vals, xref, 0.0, y);
        auto ref = densemul(pntr, indx, vals, xref);
        for (auto &r : ref) r *= 2.0;
        CHECK(diffsum(y, ref) == 0.0);
        auto mat = hala::make_sparse_matrix(mengine, 5, pntr, indx, vals);
        mat.gemv('T', 1.0, xref, 0.0, y);
        CHECK(diffsum(y, densemul(pntr, indx, vals, xref, true)) == 0.0);
        // triangular solve and the ILU known answer (tests/sparse_tests.hpp:259-311, :423-440)
        {
            std::vector<int> tp = {0, 1, 3, 5}, ti = {0, 0, 1, 0, 2};
            std::vector<double> tv = {1.0, 4.0, 2.0, 5.0, 3.0}, bN = {-6.0, -32.0, -36.0}, tx;
            auto tri = hala::make_triangular_matrix(mengine, 'L', 'N', tp, ti, tv);
            hala::sparse_trsv('N', tri, 0.5, bN, tx);
            CHECK(diffsum(tx, {-3.0, -2.0, -1.0}) < 1e-14);
            std::vector<int> ip = {0, 3, 6, 9}, ii = {0, 1, 2, 0, 1, 2, 0, 1, 2};
            std::vector<double> iv = {3.0, 2.0, 1.0, 2.0, 4.0, 3.0, 2.0, 1.0, 6.0}, r;
            auto milu = hala::make_ilu(mengine, ip, ii, iv);
            hala::ilu_apply(milu, std::vector<double>{1.0, 2.0, 3.0}, r);
            CHECK(diffsum(r, {1.0 / 9.0, 1.0 / 9.0, 4.0 / 9.0}) < 1e-14);
        }
        // solvers: the device result copied back into the std::vector
        auto b = densemul(pntr, indx, vals, xref);
        auto ilu = hala::make_ilu(mengine, pntr, indx, vals);
        hala::solve_cg(mengine, 2, pntr, indx, vals,
                       [&](hala::gpu_vector<double> const &inx, hala::gpu_vector<double> &outr) {
                           hala::ilu_apply(ilu.cilu(), inx, outr);
                       },
                       b, x);
        CHECK(x.size() == 5 && diffsum(x, xref) < 1e-14);
        x.clear();
        hala::solve_cg_ilu(mengine, 1.E-9, pntr, indx, vals, ilu, b, x);
        CHECK(diffsum(x, xref) < 1e-12);
        std::vector<double> xg;
        auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
        hala::gpu_vector<double> gx;
        hala::solve_gmres_ilu(engine, 1.E-10, 5, gp, gi, gv, hala::make_gpu_vector(b), gx);
        gx.unload(xg);
        x.clear();
        hala::solve_gmres_ilu(mengine, 1.E-10, 5, pntr, indx, vals, b, x);  // factorisation behind the scene
        CHECK(diffsum(x, xg) == 0.0);
        x.clear();
        hala::solve_gmres(mengine, 1.E-10, 5, pntr, indx, vals,
                          [&](hala::gpu_vector<double> const &inx, hala::gpu_vector<double> &outr) {
                              hala::vcopy(engine, inx, outr);
                          },
                          b, x);
        CHECK(diffsum(x, xref) < 1e-9);
        std::vector<double> B, X;
        for (int k = 0; k < 3; k++) B.insert(B.end(), b.begin(), b.end());
        hala::solve_batch_cg_ilu(mengine, {1.E-9, 6}, pntr, indx, vals, ilu, B, X);
        CHECK(X.size() == 15 && diffsum(std::vector<double>(X.begin() + 10, X.end()), xref) < 1e-12);
        std::printf("PASS mixed_engine\n");
    }
    // the four scalar types of tests/solvers_tests.cpp:14-40
    typed_trsm<double>(engine, "double", 1e-12);
    typed_trsm<float>(engine, "float", 1e-4);
    typed_trsm<std::complex<double>>(engine, "complex<double>", 1e-12);
    typed_trsm<std::complex<float>>(engine, "complex<float>", 1e-4);
    typed_solvers<double>(engine, "double", 1e-13);
    typed_solvers<float>(engine, "float", 1e-5);
    typed_solvers<std::complex<double>>(engine, "complex<double>", 1e-13);
    typed_solvers<std::complex<float>>(engine, "complex<float>", 1e-5);
    // tests/solvers_tests.hpp:105-137 (GMRES)
    {
        std::vector<int> pntr = {0, 2, 5, 8, 11, 12};
        std::vector<int> indx = {0, 1, 0, 1, 3, 1, 2, 3, 0, 2, 3, 4};
        std::vector<double> vals = {2.0, 1.0, 1.0, 3.0, 1.0, 1.0, 4.0, 1.0, 1.0, 1.0, 5.0, 2.0};
        std::vector<double> xref = {1.0, 2.0, 3.0, 4.0, 5.0}, x, x2;
        auto b = densemul(pntr, indx, vals, xref);
        auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
        auto gb = hala::make_gpu_vector(b);
        hala::gpu_vector<double> gx, gx2;
        hala::solve_gmres_ilu(engine, 1.E-9, 6, gp, gi, gv, gb, gx);
        gx.unload(x);
        CHECK(diffsum(x, xref) < 1e-8);
        // the host-driven form with a user preconditioner computes the same iterate
        auto ilu = hala::make_ilu(engine, gp, gi, gv, 'N');
        hala::solve_gmres(engine, 1.E-9, 6, gp, gi, gv,
                          [&](hala::gpu_vector<double> const &inx, hala::gpu_vector<double> &outr) {
                              hala::ilu_apply(ilu, inx, outr);
                          },
                          gb, gx2);
        gx2.unload(x2);
        CHECK(diffsum(x2, x) == 0.0);
        std::printf("PASS solve_gmres\n");
    }
    // examples/example_gpu.cpp:90-98 on the GPU engine, and GPU == CPU-oracle agreement (:363-365)
    {
        int sides[6] = {4, 8, 16, 32, 64, 128};
        int iters[6] = {9, 14, 23, 41, 78, 152};
        for (int t = 0; t < 6; t++) {
            int n = 0;
            int64_t nnz = 0;
            hala::check_hala(hala_gen_discretize_pde(sides[t], &n, &nnz, nullptr, nullptr, nullptr, nullptr), "gen");
            std::valarray<int> pntr(n + 1), indx((size_t)nnz);
            std::valarray<double> vals((size_t)nnz), f(n), u(0.0, n);
            hala::check_hala(hala_gen_discretize_pde(sides[t], &n, &nnz, std::begin(pntr), std::begin(indx),
                                                     std::begin(vals), std::begin(f)),
                             "gen");
            auto gp = engine.load(pntr);
            auto gi = engine.load(indx);
            auto gv = engine.load(vals);
            auto gf = engine.load(f);
            auto gu = engine.load(u);
            auto ilu = hala::make_ilu(engine, gp, gi, gv, 'N');
            int it = hala::solve_cg(engine, 1.E-8, gp, gi, gv,
                                    [&](hala::gpu_vector<double> const &inx, hala::gpu_vector<double> &outr) {
                                        hala::ilu_apply(ilu, inx, outr);
                                    },
                                    gf, gu);
            CHECK(it == iters[t]);
            auto gu2 = engine.load(std::valarray<double>(0.0, n));
            int it2 = hala::solve_cg_ilu(engine, 1.E-8, gp, gi, gv, ilu, gf, gu2);
            CHECK(it2 == it);
            hala::axpy(engine, -1.0, gu, gu2);
            CHECK(hala::norm2(engine, gu2) == 0.0);  // fused device path == host-driven loop, bit for bit
        }
        std::printf("PASS example_gpu table\n");
    }
    // the same table through a user PCG written in the no-engine style, bound to the
    // gpu_engine: equal iteration counts, and bit for bit the device-resident solve
    {
        int sides[6] = {4, 8, 16, 32, 64, 128};
        int iters[6] = {9, 14, 23, 41, 78, 152};
        for (int t = 0; t < 6; t++) {
            int n = 0;
            int64_t nnz = 0;
            hala::check_hala(hala_gen_discretize_pde(sides[t], &n, &nnz, nullptr, nullptr, nullptr, nullptr), "gen");
            std::valarray<int> pntr(n + 1), indx((size_t)nnz);
            std::valarray<double> vals((size_t)nnz), f(n), u(0.0, n);
            hala::check_hala(hala_gen_discretize_pde(sides[t], &n, &nnz, std::begin(pntr), std::begin(indx),
                                                     std::begin(vals), std::begin(f)),
                             "gen");
            auto gp = engine.load(pntr);
            auto gi = engine.load(indx);
            auto gv = engine.load(vals);
            auto gf = engine.load(f);
            auto gu = engine.load(u);
            int it = user_pcg(engine, 100000, 1.E-8, gp, gi, gv, gf, gu);
            CHECK(it == iters[t]);
            auto gu2 = engine.load(u);
            CHECK(hala::solve_cg_ilu(engine, 1.E-8, gp, gi, gv, gf, gu2) == it);
            hala::axpy(engine, -1.0, gu, gu2);
            CHECK(hala::norm2(engine, gu2) == 0.0);
        }
        std::printf("PASS no-engine user PCG (example_gpu table)\n");
    }
    // the no-engine solver forms on bound vectors (hex/solvers/hala_solvers_cg.hpp:181-227,
    // :295-352), and float callers passing stop_criteria<float>
    {
        std::vector<int> pntr = {0, 3, 6, 9, 12, 15};
        std::vector<int> indx = {0, 1, 4, 0, 1, 2, 1, 2, 3, 2, 3, 4, 0, 3, 4};
        std::vector<double> vals = {4.0, 1.0, 1.0, 1.0, 4.0, 1.0, 1.0, 4.0, 1.0, 1.0, 4.0, 1.0, 1.0, 1.0, 4.0};
        std::vector<double> xref = {1.0, -2.0, 3.0, -4.0, 5.0};
        auto b = densemul(pntr, indx, vals, xref);
        auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
        auto gb = hala::make_gpu_vector(b);
        hala::gpu_vector<double> gx1, gx2, gx3, gx4, gx5;
        auto bp = hala::bind_engine_vector(engine, gp);
        auto bi = hala::bind_engine_vector(engine, gi);
        auto bv = hala::bind_engine_vector(engine, gv);
        auto bb = hala::bind_engine_vector(engine, gb);
        CHECK(hala::check_engines(bp, bi, bv, bb));
        auto ilu = hala::make_ilu(bp, bi, bv);
        std::vector<double> x;
        {
            auto bx = hala::bind_engine_vector(engine, gx1);
            int it = hala::solve_cg_ilu({1.E-12, 50}, bp, bi, bv, ilu, bb, bx);
            CHECK(it > 1 && it < 50);
        }
        gx1.unload(x);
        CHECK(diffsum(x, xref) < 1e-10);
        {  // preconditioner_noe: the lambda receives bound vectors
            auto bx = hala::bind_engine_vector(engine, gx2);
            hala::solve_cg(hala::stop_criteria<double>(1.E-12, 50), bp, bi, bv,
                           [&](auto const &in, auto &out) { hala::ilu_apply(ilu, in, out); }, bb, bx);
        }
        hala::axpy(engine, -1.0, gx1, gx2);
        CHECK(hala::norm2(engine, gx2) == 0.0);  // host loop == device-resident loop, bit for bit
        {
            auto bx = hala::bind_engine_vector(engine, gx3);
            hala::solve_cg_ilu(1.E-12, bp, bi, bv, bb, bx);  // factor on the fly
            auto bx4 = hala::bind_engine_vector(engine, gx4);
            hala::solve_gmres_ilu(1.E-12, 5, bp, bi, bv, ilu, bb, bx4);
            auto bx5 = hala::bind_engine_vector(engine, gx5);
            hala::solve_gmres(1.E-12, 5, bp, bi, bv, [&](auto const &in, auto &out) { hala::ilu_apply(ilu, in, out); },
                              bb, bx5);
        }
        gx3.unload(x);
        CHECK(diffsum(x, xref) < 1e-10);
        gx4.unload(x);
        CHECK(diffsum(x, xref) < 1e-10);
        hala::axpy(engine, -1.0, gx4, gx5);
        CHECK(hala::norm2(engine, gx5) == 0.0);
        {  // three right-hand sides, column-major
            std::vector<double> B;
            for (int k = 0; k < 3; k++) B.insert(B.end(), b.begin(), b.end());
            auto gB = hala::make_gpu_vector(B);
            hala::gpu_vector<double> gX;
            auto bB = hala::bind_engine_vector(engine, gB);
            auto bX = hala::bind_engine_vector(engine, gX);
            hala::solve_batch_cg_ilu({1.E-12, 50}, bp, bi, bv, ilu, bB, bX);
            std::vector<double> X;
            gX.unload(X);
            CHECK(X.size() == 15 && diffsum(std::vector<double>(X.begin() + 10, X.end()), xref) < 1e-10);
        }
        // float values: the tolerance is a float, compared in float
        std::vector<float> vals32(vals.begin(), vals.end()), b32(b.begin(), b.end());
        auto gv32 = hala::make_gpu_vector(vals32);
        auto gb32 = hala::make_gpu_vector(b32);
        hala::gpu_vector<float> gx32;
        auto ilu32 = hala::make_ilu(engine, gp, gi, gv32);
        static_assert(std::is_same<hala::get_precision_type<hala::gpu_vector<std::complex<float>>>, float>::value,
                      "precision of complex<float> is float");
        int it32 = hala::solve_cg_ilu(engine, hala::stop_criteria<float>(1.E-4f, 50), gp, gi, gv32, ilu32, gb32, gx32);
        std::vector<float> x32;
        gx32.unload(x32);
        double e32 = 0;
        for (int k = 0; k < 5; k++) e32 += std::fabs(x32[k] - xref[k]);
        CHECK(it32 > 1 && it32 < 50 && e32 < 1e-4);
        hala::gpu_vector<float> gy32;
        auto bv32 = hala::bind_engine_vector(engine, gv32);
        auto bb32 = hala::bind_engine_vector(engine, gb32);
        auto by32 = hala::bind_engine_vector(engine, gy32);
        CHECK(hala::solve_cg_ilu(hala::stop_criteria<float>(1.E-4f, 50), bp, bi, bv32, ilu32, bb32, by32) == it32);
        hala::axpy(engine, -1.0f, gx32, gy32);
        CHECK(hala::norm2(engine, gy32) == 0.0f);
        std::printf("PASS no-engine solvers, stop_criteria<float>\n");
    }
    // mixed precision refinement reaches the FP64 residual
    {
        int n = 0;
        int64_t nnz = 0;
        hala::check_hala(hala_gen_convdiff27(12, 3, 10.0, &n, &nnz, nullptr, nullptr, nullptr), "gen");
        std::vector<int> pntr(n + 1), indx((size_t)nnz);
        std::vector<double> vals((size_t)nnz), b(n);
        hala::check_hala(hala_gen_convdiff27(12, 3, 10.0, &n, &nnz, pntr.data(), indx.data(), vals.data()), "gen");
        hala::check_hala(hala_gen_uniform(5, n, -1.0, 1.0, b.data()), "gen");
        auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
        auto gb = hala::make_gpu_vector(b);
        hala::gpu_vector<double> gx, r;
        hala::solve_gmres_ir(engine, 1e-10, 30, gp, gi, gv, gb, gx);
        hala::vcopy(engine, gb, r);
        hala::sparse_gemv(engine, 'N', n, n, -1.0, gp, gi, gv, gx, 1.0, r);
        CHECK(hala::norm2(engine, r) < 1e-10 * hala::norm2(engine, gb));
        std::printf("PASS solve_gmres_ir\n");
    }
    // library errors surface as std::runtime_error
    {
        bool thrown = false;
        try {
            std::vector<int> pntr = {0, 2, 3}, indx = {1, 0, 5};
            std::vector<double> vals = {1, 1, 1};
            auto gp = hala::make_gpu_vector(pntr);
        auto gi = hala::make_gpu_vector(indx);
        auto gv = hala::make_gpu_vector(vals);
            auto m = hala::make_sparse_matrix(engine, 2, gp, gi, gv);
        } catch (std::runtime_error &) {
            thrown = true;
        }
        CHECK(thrown);
        std::printf("PASS errors\n");
    }
    std::printf(failures ? "FAILED %d\n" : "ALL PASSED\n", failures);
    return failures ? 1 : 0;
}
