// hala_mixed.hpp -- the mixed engine: inputs and outputs in host memory, the
// work on the GPU.  Each call loads its inputs into device vectors, runs the
// gpu_engine path and copies its outputs back (wax/hala_lib_extensions.hpp:
// 68-81 mixed_engine, :127-220 mixed_sparse_matrix, :255-293 the triangular
// matrix, :317-376 mixed_ilu; BLAS overloads wax/hala_mixed_overloads.hpp).
// Matrices and factors keep their own device copies of the pattern and values,
// so a solver loads the operator once and the vectors once per call.
// Included by hala.hpp.
#pragma once

namespace hala {

struct mixed_engine {
    explicit mixed_engine(gpu_engine const &ingpu) : cengine(ingpu) {}
    gpu_engine const &gpu() const { return cengine; }
    template <typename T>
    using dfvector = std::vector<T>;
    gpu_engine cengine;
};

// ---- vectors ----------------------------------------------------------------
template <class VectorLike>
auto new_vector(mixed_engine const &, VectorLike const &) {
    return std::vector<typename host_traits<std::remove_cv_t<VectorLike>>::value_type>();
}
template <class VectorLike>
void set_zero(mixed_engine const &, size_t num_entries, VectorLike &&x) {
    using H = host_traits<std::remove_cv_t<std::remove_reference_t<VectorLike>>>;
    if (H::size(x) < num_entries) H::resize(x, num_entries);
    std::fill(H::data(x), H::data(x) + num_entries, typename H::value_type(0));
}

// ---- BLAS-1 (wax/hala_mixed_overloads.hpp: vcopy, dot, norm2, axpy, scal) ------
template <class VX, class VY>
void vcopy(mixed_engine const &e, VX const &x, VY &&y) {
    auto gx = e.gpu().load(x);
    auto gy = gpu_bind_vector(e.gpu(), y);
    vcopy(e.gpu(), gx, gy);
}
template <class VX, class VY>
auto dot(mixed_engine const &e, VX const &x, VY const &y) {
    return dot(e.gpu(), e.gpu().load(x), e.gpu().load(y));
}
template <class VX>
auto norm2(mixed_engine const &e, VX const &x) {
    return norm2(e.gpu(), e.gpu().load(x));
}
template <typename FP, class VX, class VY>
void axpy(mixed_engine const &e, FP alpha, VX const &x, VY &&y) {
    auto gx = e.gpu().load(x);
    auto gy = gpu_bind_vector(e.gpu(), y);
    axpy(e.gpu(), alpha, gx, gy);
}
template <typename FP, class VX>
void scal(mixed_engine const &e, FP alpha, VX &&x) {
    auto gx = gpu_bind_vector(e.gpu(), x);
    scal(e.gpu(), alpha, gx);
}

// ---- CSR matrix (wax/hala_lib_extensions.hpp:127-237) ---------------------------
template <typename T>
struct mixed_sparse_matrix {
    using value_type = T;
    using engine_type = mixed_engine;
    template <class VP, class VI, class VV>
    mixed_sparse_matrix(mixed_engine const &engine, int num_rows, int num_cols, int num_nz, VP const &pntr,
                        VI const &indx, VV const &vals)
        : rpntr(engine.gpu().load(pntr, num_rows + 1)), rindx(engine.gpu().load(indx, num_nz)),
          rvals(engine.gpu().load(vals, num_nz)),
          mat(make_sparse_matrix(engine.gpu(), num_rows, num_cols, num_nz, rpntr, rindx, rvals)) {}
    template <class VP, class VI, class VV>
    mixed_sparse_matrix(mixed_engine const &engine, int num_cols, VP const &pntr, VI const &indx, VV const &vals)
        : rpntr(engine.gpu().load(pntr)), rindx(engine.gpu().load(indx)), rvals(engine.gpu().load(vals)),
          mat(make_sparse_matrix(engine.gpu(), num_cols, rpntr, rindx, rvals)) {}
    mixed_sparse_matrix(mixed_sparse_matrix const &) = delete;
    mixed_sparse_matrix(mixed_sparse_matrix &&) = default;
    mixed_sparse_matrix &operator=(mixed_sparse_matrix const &) = delete;
    mixed_sparse_matrix &operator=(mixed_sparse_matrix &&) = default;

    template <typename FPa, class VX, typename FPb, class VY>
    size_t gemv_buffer_size(char, FPa, VX const &, FPb, VY &&) const {
        return 0;
    }
    template <typename FPa, class VX, typename FPb, class VY>
    void gemv(char trans, FPa alpha, VX const &x, FPb beta, VY &&y) const {
        auto gx = mat.engine().load(x);
        auto gy = gpu_bind_vector(mat.engine(), y);
        mat.gemv(trans, alpha, gx, beta, gy);
    }
    template <typename FPa, class VX, typename FPb, class VY, class VB>
    void gemv(char trans, FPa alpha, VX const &x, FPb beta, VY &&y, VB &&) const {
        gemv(trans, alpha, x, beta, y);
    }
    template <typename FSA, class VB, typename FSB, class VC>
    size_t gemm_buffer_size(char, char, int, int, FSA, VB const &, int, FSB, VC &&, int) const {
        return 0;
    }
    template <typename FSA, class VB, typename FSB, class VC>
    void gemm(char transa, char transb, int b_rows, int b_cols, FSA alpha, VB const &B, int ldb, FSB beta, VC &&C,
              int ldc) const {
        const int N = (transb == 'N' || transb == 'n') ? b_cols : b_rows;
        if ((size_t)C.size() < (size_t)ldc * (size_t)N) C.resize((size_t)ldc * (size_t)N);
        auto gB = mat.engine().load(B);
        auto gC = gpu_bind_vector(mat.engine(), C);
        mat.gemm(transa, transb, b_rows, b_cols, alpha, gB, ldb, beta, gC, ldc);
    }
    template <typename FSA, class VB, typename FSB, class VC, class VT>
    void gemm(char transa, char transb, int b_rows, int b_cols, FSA alpha, VB const &B, int ldb, FSB beta, VC &&C,
              int ldc, VT &&) const {
        gemm(transa, transb, b_rows, b_cols, alpha, B, ldb, beta, C, ldc);
    }

    gpu_vector<int> rpntr, rindx;
    gpu_vector<T> rvals;
    gpu_sparse_matrix<T> mat;
};

template <class VP, class VI, class VV>
auto make_sparse_matrix(mixed_engine const &engine, int num_rows, int num_cols, int num_nz, VP const &pntr,
                        VI const &indx, VV const &vals) {
    return mixed_sparse_matrix<scalar_of<VV>>(engine, num_rows, num_cols, num_nz, pntr, indx, vals);
}
template <class VP, class VI, class VV>
auto make_sparse_matrix(mixed_engine const &engine, int num_cols, VP const &pntr, VI const &indx, VV const &vals) {
    return mixed_sparse_matrix<scalar_of<VV>>(engine, num_cols, pntr, indx, vals);
}
template <typename FPa, class VP, class VI, class VV, class VX, typename FPb, class VY>
void sparse_gemv(mixed_engine const &engine, char trans, int M, int N, FPa alpha, VP const &pntr, VI const &indx,
                 VV const &vals, VX const &x, FPb beta, VY &&y) {
    auto gp = engine.gpu().load(pntr), gi = engine.gpu().load(indx);
    auto gv = engine.gpu().load(vals), gx = engine.gpu().load(x);
    auto gy = gpu_bind_vector(engine.gpu(), y);
    sparse_gemv(engine.gpu(), trans, M, N, alpha, gp, gi, gv, gx, beta, gy);
}

// ---- triangular matrix (wax/hala_lib_extensions.hpp:255-301) -------------------
template <typename T>
struct mixed_triangular_matrix {
    using value_type = T;
    using engine_type = mixed_engine;
    template <class VP, class VI, class VV>
    mixed_triangular_matrix(mixed_engine const &eng, char uplo, char diag, VP const &pn, VI const &in, VV const &va,
                            char policy)
        : gpu_pntr(eng.gpu().load(pn)), gpu_indx(eng.gpu().load(in)), gpu_vals(eng.gpu().load(va)),
          tri(eng.gpu(), uplo, diag, gpu_pntr, gpu_indx, gpu_vals, policy), rengine(eng) {}
    mixed_triangular_matrix(mixed_triangular_matrix const &) = delete;
    mixed_triangular_matrix(mixed_triangular_matrix &&) = default;
    mixed_triangular_matrix &operator=(mixed_triangular_matrix const &) = delete;
    mixed_triangular_matrix &operator=(mixed_triangular_matrix &&) = default;

    gpu_triangular_matrix<T> const &gpu_tri() const { return tri; }
    template <typename FP, class VB, class VX>
    void trsv(char trans, FP alpha, VB const &b, VX &&x) const {
        auto gb = rengine.gpu().load(b);
        auto gx = gpu_bind_vector(rengine.gpu(), x);
        tri.trsv(trans, alpha, gb, gx);
    }
    template <typename FP, class VB, class VX, class VW>
    void trsv(char trans, FP alpha, VB const &b, VX &&x, VW &&) const {
        trsv(trans, alpha, b, x);
    }
    template <typename FP, class VB>
    void trsm(char transa, char transb, int nrhs, FP alpha, VB &&B, int ldb = -1) const {
        auto gB = gpu_bind_vector(rengine.gpu(), B);
        tri.trsm(transa, transb, nrhs, alpha, gB, ldb);
    }
    template <typename FP, class VB, class VW>
    void trsm(char transa, char transb, int nrhs, FP alpha, VB &&B, int ldb, VW &&) const {
        trsm(transa, transb, nrhs, alpha, B, ldb);
    }

    gpu_vector<int> gpu_pntr, gpu_indx;
    gpu_vector<T> gpu_vals;
    gpu_triangular_matrix<T> tri;
    mixed_engine rengine;
};
template <class VP, class VI, class VV>
auto make_triangular_matrix(mixed_engine const &engine, char uplo, char diag, VP const &pntr, VI const &indx,
                            VV const &vals, char policy = 'N') {
    return mixed_triangular_matrix<scalar_of<VV>>(engine, uplo, diag, pntr, indx, vals, policy);
}

// ---- ILU(0) (wax/hala_lib_extensions.hpp:317-376) ------------------------------
template <typename T>
struct mixed_ilu {
    using value_type = T;
    using engine_type = mixed_engine;
    template <class VP, class VI, class VV>
    mixed_ilu(mixed_engine const &eng, VP const &mpntr, VI const &mindx, VV const &mvals, char policy)
        : rengine(eng), num_rows((int)host_traits<VP>::size(mpntr) - 1), gpu_pntr(eng.gpu().load(mpntr)),
          gpu_indx(eng.gpu().load(mindx)), gilu(eng.gpu(), gpu_pntr, gpu_indx, eng.gpu().load(mvals), policy) {}
    mixed_ilu(mixed_ilu const &) = delete;
    mixed_ilu(mixed_ilu &&) = default;
    mixed_ilu &operator=(mixed_ilu const &) = delete;
    mixed_ilu &operator=(mixed_ilu &&) = default;

    // r = (LU)^-1 x, num_rhs columns of num_rows entries in host memory
    template <class VX, class VR>
    void apply(VX const &x, VR &&r, int num_rhs = 1) const {
        using H = host_traits<std::remove_cv_t<std::remove_reference_t<VR>>>;
        if (H::size(r) < (size_t)num_rows * num_rhs) H::resize(r, (size_t)num_rows * num_rhs);
        auto gx = rengine.gpu().load(x);
        gpu_vector<T> gr((size_t)num_rows * num_rhs, rengine.gpu().device());
        gilu.apply(gx, gr, num_rhs);
        gr.unload(r);
    }
    template <class VX, class VR, class VT>
    void apply(VX const &x, VR &&r, int num_rhs, VT &&) const {
        apply(x, r, num_rhs);
    }
    mixed_engine const &engine() const { return rengine; }
    gpu_ilu<T> const &cilu() const { return gilu; }

  private:
    mixed_engine rengine;
    int num_rows;
    gpu_vector<int> gpu_pntr, gpu_indx;  // the factor aliases the pattern
    gpu_ilu<T> gilu;
};
template <class VP, class VI, class VV>
auto make_ilu(mixed_engine const &engine, VP const &pntr, VI const &indx, VV const &vals, char policy = 'N') {
    return mixed_ilu<scalar_of<VV>>(engine, pntr, indx, vals, policy);
}

}  // namespace hala
