// blas1.hpp -- canonical reductions shared by the solvers.
#pragma once
#include "common.hpp"

namespace hala_amd {
// out_dev[0] = cdot(x, y) (mode 0), cdot(x, x) (mode 1) or sqrt(cdot(x, x)) (mode 2),
// canonical order; asynchronous on the context stream.
template <class T>
void cdot_launch(hala_ctx c, int64_t n, const T *x, const T *y, T *out_dev, int mode);
}  // namespace hala_amd
