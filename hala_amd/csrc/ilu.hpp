// ilu.hpp -- ILU(0) handle (the gpu_ilu of gpu/hala_gpu_ilu.hpp:45-195).
#pragma once
#include <functional>
#include <memory>
#include "spmv.hpp"
#include "trsolve.hpp"
#include "trstencil.hpp"

struct hala_ilu_s {
    hala_ctx ctx = nullptr;
    int n = 0;
    int64_t nnz = 0;
    const int *pntr = nullptr;  // aliased, as the reference's gpu_ilu
    const int *indx = nullptr;
    hala_amd::dbuf diag;       // diag[i]: position of a_ii
    hala_amd::dbuf factor;     // combined L\U values (CSR order), FP64
    hala_amd::dbuf dpos, pptr, pairs;  // elimination: u_ii position and (ik, jk) pairs per lower entry
    hala_amd::dbuf frow, fnp;          // rows in level order: (start, end, diagonal, first pair), pair count
    hala_amd::dbuf fwork;              // in-progress values of rows too long to stage in LDS (factor indexing)
    hala_amd::tri_plan L, U;   // forward (unit L) and backward (U) solves
    hala_amd::tri_plan Ut, Lt; // adjoint: U^T forward, L^T backward (built on demand)
    hala_amd::csr_transpose tp;
    hala_amd::dbuf tmp64, tmp32;  // natural-order hand-off between the two solves
    bool factored = false, has32 = false, adjoint = false, adjoint_loaded = false;
    // Jacobi-applied factor (ilujac.hip): packed strict-L / strict-U parts with
    // FP32 (and, for FP64 applies, FP64) values and diagonal, four iterates of
    // each precision (built on first use)
    hala_amd::dbuf jl_ptr, jl_idx, jl_val, ju_ptr, ju_idx, ju_val, jd_val, jwork;
    hala_amd::dbuf jl_off, ju_off;  // the strict parts' columns as int16 block offsets (empty: not banded)
    hala_amd::dbuf jl_val64, ju_val64, jd_val64, jwork64;
    bool jac_pattern = false, jac_values = false, jac_values64 = false;
    // complex factors (cplx.hip): interleaved real/imaginary, CSR order
    hala_amd::dbuf zfactor, cfactor;
    // pencil schedule of the two sweeps when the pattern is a stencil (trstencil.hip), else null
    std::unique_ptr<hala_amd::stencil_plan> sp;
    bool zfactored = false, cfactored = false;
};

struct hala_tri_s {
    hala_ctx ctx = nullptr;
    int n = 0;
    int64_t nnz = 0;
    const int *pntr = nullptr;
    const int *indx = nullptr;
    char uplo = 'L', diag = 'N';
    hala_amd::tri_plan N, T;  // plain and transposed solves (T built on demand)
    hala_amd::csr_transpose tp;
    bool t_built = false;
};

namespace hala_amd {
// Multi-right-hand-side triangular solve (hala_trsm_*): B <- op(T)^-1 (alpha B)
// column by column through `col`, which solves one column from a contiguous
// input into a contiguous output.  esize: bytes per element.
void tri_trsm_cols(hala_tri T, char transb, int nrhs, size_t esize, void *B, int ldb,
                   const std::function<void(const void *, void *)> &col);
// r = (LU)^-1 x in the requested precision (see hala_ilu_apply_*)
// gate (nullable, device flag): skip the apply when *gate != 0 (solver stop flag)
void ilu_apply64(hala_ilu M, const double *x, double *r, const int *gate = nullptr);
// Both triangles on the column layout: the apply's ends can be fused by the
// caller.  ilu_apply64_mid runs L from its gathered input (L's col.xs64 filled
// by the caller; the polled outputs were re-armed by the previous apply's
// sweeps), which hands its results to U's input itself, then U; the result
// stays in U.out64 in slot order.  ax (nullable): a vector update the L
// sweep's helper workgroups do meanwhile (see col_sweep64).
bool ilu_fused64(hala_ilu M);
void ilu_apply64_mid(hala_ilu M, const int *gate, const hala_amd::col_axpy *ax = nullptr);
void ilu_apply32x64(hala_ilu M, const double *x, double *r, const int *gate = nullptr);
void ilu_apply_t64(hala_ilu M, const double *x, double *r, const int *gate = nullptr);
void ilu_prepare_adjoint(hala_ilu M);
// r = (LU)^-1 x with the FP32 factor applied by k Jacobi sweeps per triangle (ilujac.hip)
void ilu_apply32x64_jacobi(hala_ilu M, const double *x, double *r, const int *gate, int k);
// r = (LU)^-1 x with the FP64 factor applied by k Jacobi sweeps per triangle (ilujac.hip)
void ilu_apply64_jacobi(hala_ilu M, const double *x, double *r, const int *gate, int k);
}  // namespace hala_amd
