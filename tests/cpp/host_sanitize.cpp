// Host code of libhala_amd under AddressSanitizer / UndefinedBehaviorSanitizer
// (built by tests/test_cpu_sanitize.py with g++, no GPU code): the Matrix
// Market reader and writer on the files named on the command line, and the
// synthetic generators on small and degenerate sizes, every buffer sized
// exactly from the sizing call.  Prints one status line per file; any memory
// or undefined-behaviour error aborts with the sanitizer's report.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "hala_amd.h"

namespace hala_amd {
// the library's error slot (defined in ctx.hip for the real build)
static std::string last;
void set_error(const std::string &msg) { last = msg; }
}  // namespace hala_amd

static int roundtrip(const char *path, const char *tmp) {
    int nr = 0, nc = 0, field = 0, sym = 0;
    int64_t stored = 0, expanded = 0;
    int st = hala_mm_info(path, &nr, &nc, &stored, &expanded, &field, &sym);
    if (st != HALA_OK) return st;
    if (expanded > (1 << 22)) return -100;  // a mutated size line: not read
    for (int expand = 0; expand < 2; expand++) {
        const int64_t cap = expand ? expanded : stored;
        std::vector<int> rows((size_t)cap + 1), cols((size_t)cap + 1);
        std::vector<double> vals((size_t)cap + 1);
        int64_t cnt = 0;
        st = hala_mm_read_coo(path, expand, cap, rows.data(), cols.data(), vals.data(), &cnt);
        if (st != HALA_OK) return st;
        // CSR by counting, then written and read back
        std::vector<int> pntr((size_t)nr + 1, 0), indx((size_t)cnt + 1);
        std::vector<double> v((size_t)cnt + 1);
        for (int64_t e = 0; e < cnt; e++) pntr[(size_t)rows[e] + 1]++;
        for (int r = 0; r < nr; r++) pntr[r + 1] += pntr[r];
        std::vector<int> fill(pntr.begin(), pntr.end() - 1);
        for (int64_t e = 0; e < cnt; e++) {
            const int at = fill[rows[e]]++;
            indx[at] = cols[e];
            v[at] = vals[e];
        }
        st = hala_mm_write_csr(tmp, nr, nc, pntr.data(), indx.data(), v.data());
        if (st != HALA_OK) return st;
        int nr2 = 0, nc2 = 0, f2 = 0, s2 = 0;
        int64_t st2 = 0, ex2 = 0;
        st = hala_mm_info(tmp, &nr2, &nc2, &st2, &ex2, &f2, &s2);
        if (st != HALA_OK || nr2 != nr || nc2 != nc || st2 != cnt) return -101;
    }
    return HALA_OK;
}

static bool generators() {
    bool ok = true;
    for (int a = 1; a <= 6; a++)
        for (int b = 1; b <= 5; b++) {
            int n = 0;
            int64_t z = 0;
            if (hala_gen_laplace2d(a, b, 4.0, -1.0, &n, &z, nullptr, nullptr, nullptr) == HALA_OK) {
                std::vector<int> p((size_t)n + 1), i((size_t)z);
                std::vector<double> v((size_t)z);
                ok = ok && hala_gen_laplace2d(a, b, 4.0, -1.0, &n, &z, p.data(), i.data(), v.data()) == HALA_OK &&
                     p[n] == z;
            }
            for (int c = 1; c <= 4; c++)
                if (hala_gen_laplace3d(a, b, c, &n, &z, nullptr, nullptr, nullptr) == HALA_OK) {
                    std::vector<int> p((size_t)n + 1), i((size_t)z);
                    std::vector<double> v((size_t)z);
                    ok = ok && hala_gen_laplace3d(a, b, c, &n, &z, p.data(), i.data(), v.data()) == HALA_OK &&
                         p[n] == z;
                }
        }
    for (int m = 1; m <= 6; m++) {
        int n = 0;
        int64_t z = 0;
        if (hala_gen_convdiff27(m, 3 + m, 10.0, &n, &z, nullptr, nullptr, nullptr) == HALA_OK) {
            std::vector<int> p((size_t)n + 1), i((size_t)z);
            std::vector<double> v((size_t)z);
            ok = ok && hala_gen_convdiff27(m, 3 + m, 10.0, &n, &z, p.data(), i.data(), v.data()) == HALA_OK &&
                 p[n] == z;
        }
        if (hala_gen_discretize_pde(m, &n, &z, nullptr, nullptr, nullptr, nullptr) == HALA_OK) {
            std::vector<int> p((size_t)n + 1), i((size_t)z);
            std::vector<double> v((size_t)z), f((size_t)n);
            ok = ok && hala_gen_discretize_pde(m, &n, &z, p.data(), i.data(), v.data(), f.data()) == HALA_OK &&
                 p[n] == z;
        }
    }
    // invalid sizes are refused, not overflowed
    int n = 0;
    int64_t z = 0;
    ok = ok && hala_gen_laplace3d(0, 4, 4, &n, &z, nullptr, nullptr, nullptr) != HALA_OK;
    ok = ok && hala_gen_laplace3d(2000, 2000, 2000, &n, &z, nullptr, nullptr, nullptr) != HALA_OK;
    ok = ok && hala_mm_write_csr("/nonexistent-dir/x.mtx", -1, 3, &n, nullptr, nullptr) != HALA_OK;
    return ok;
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const char *tmp = argv[1];
    for (int a = 2; a < argc; a++) std::printf("%s %d\n", argv[a], roundtrip(argv[a], tmp));
    std::printf("generators %s\n", generators() ? "ok" : "FAILED");
    return 0;
}
