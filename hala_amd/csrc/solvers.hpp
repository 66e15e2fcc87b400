// solvers.hpp -- device-resident Krylov drivers shared state.
#pragma once
#include "ilu.hpp"

namespace hala_amd {

// Solver scalars kept on the device (inside hala_ctx_s::scal at S_SOLVER).
struct cg_state {
    double zr, alpha, beta, rn, tol;
    double *hist;
    int it, done, max_iter, max_hist, status;
};

inline cg_state *cg_state_ptr(hala_ctx c) { return reinterpret_cast<cg_state *>(c->scal.as<double>() + S_SOLVER); }

// Reads (it, done, status) back to the host; synchronises the stream.
void cg_read(hala_ctx c, int *it, int *done, int *status, double *rn);

// Work space (doubles) of one CG run over n rows: r, p, Ap, z and two sets of
// canonical-block reduction partials.
inline size_t cg_noalloc_doubles(int64_t n) { return (size_t)(4 * n + 2 * ((n + CB - 1) / CB)); }

// hala_cg_f64's solve over the caller's work space of cg_noalloc_doubles(n)
// doubles; allocates nothing (cg.hip).
void cg_noalloc(hala_csr A, const double *vals, int prec, hala_ilu ilu, const double *pdiag, const double *b,
                double *x, double tol, int max_iter, int *iterations, double *hist, int max_hist, double *work);

// The allocation-free entry point's device scratch on the current device,
// looked up once per device at context creation (noalloc.hip).
double *noalloc_scratch(int device);

}  // namespace hala_amd
