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

}  // namespace hala_amd
