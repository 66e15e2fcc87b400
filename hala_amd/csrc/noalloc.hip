// noalloc.hip -- the allocation-free CG entry point, hala_cg_f64_noalloc.
//
// hala_cg_f64 takes its work vectors (r, p, Ap, z) and reduction partials from
// buffers the context grows with hipMalloc on first use.  Callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc) use this variant instead: the same kernels and the same
// arithmetic, over a file-scope device array.  Every call on one GPU shares
// that GPU's array; calls that overlap on one GPU race.
#include <mutex>

#include "solvers.hpp"

namespace hala_amd {

// The largest system the scratch holds: 2^21 rows (the 128^3 7-point operator).
constexpr int64_t NOALLOC_MAX_ROWS = int64_t(1) << 21;
// 4 n work-vector doubles plus 2 ceil(n/256) partials: 64.06 MiB per GPU,
// held from the first use of the library on that GPU until the process ends.
constexpr size_t NOALLOC_DOUBLES = 4 * (size_t)NOALLOC_MAX_ROWS + 2 * (size_t)(NOALLOC_MAX_ROWS / CB);

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

// Device address of `scratch` per device.  In host code the name `scratch` is
// only a host-side stand-in: every kernel argument, fill or copy uses the
// address looked up here.
static std::mutex scratch_mu;
static std::vector<double *> scratch_ptr;

static std::string scratch_err;

// Called with `device` current by hala_ctx_create (the library's set-up), once
// per device; later calls return the recorded address.  A failed lookup leaves
// the rest of the library usable: it returns null, and the entry point reports
// the lookup's error.
double *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> g(scratch_mu);
    if ((int)scratch_ptr.size() <= device) scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!scratch_ptr[device]) {
        double *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();  // clear it: later calls must not report this lookup
            scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        scratch_ptr[device] = p;
    }
    return scratch_ptr[device];
}

}  // namespace hala_amd

using namespace hala_amd;

extern "C" {

int hala_cg_f64_noalloc(hala_csr A, const double *vals, int prec, hala_ilu ilu, const double *pdiag,
                        const double *b, double *x, double tol, int max_iter, int *iterations, double *hist,
                        int max_hist) {
    HALA_RANGE();
    return guard([&] {
        HALA_REQUIRE(A && vals && b && x, HALA_ERR_ARG, "hala_cg_f64_noalloc: null argument");
        HALA_REQUIRE(prec >= 0 && prec <= 2, HALA_ERR_ARG, "hala_cg_f64_noalloc: prec must be 0, 1 or 2");
        HALA_REQUIRE(!hist || max_hist >= 0, HALA_ERR_ARG, "hala_cg_f64_noalloc: bad history size");
        HALA_REQUIRE(A->nrows <= NOALLOC_MAX_ROWS, HALA_ERR_ARG,
                     "hala_cg_f64_noalloc: " + std::to_string(A->nrows) + " rows exceed the scratch of " +
                         std::to_string(NOALLOC_MAX_ROWS) + " rows (use hala_cg_f64)");
        HALA_REQUIRE(cg_noalloc_doubles(A->nrows) <= NOALLOC_DOUBLES, HALA_ERR_INTERNAL,
                     "hala_cg_f64_noalloc: scratch layout");
        hala_ctx c = A->ctx;
        set_device(c);
        if (!c->noalloc) {
            std::lock_guard<std::mutex> g(scratch_mu);
            throw status_error{HALA_ERR_HIP, "hala_cg_f64_noalloc: no scratch on this device (" + scratch_err + ")"};
        }
        cg_noalloc(A, vals, prec, ilu, pdiag, b, x, tol, max_iter, iterations, hist, hist ? max_hist : 0,
                   c->noalloc);
    });
}

}  // extern "C"
