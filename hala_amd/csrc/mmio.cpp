// mmio.cpp -- Matrix Market files (host code): the coordinate and array
// formats of the NIST exchange format ("%%MatrixMarket matrix <format>
// <field> <symmetry>", 1-based indices), read into COO triplets and written
// from CSR.  Fields real, double, integer and pattern (values 1); symmetries
// general, symmetric and skew-symmetric (the mirrored entry (j,i) follows its
// stored (i,j) when expanded, negated for skew-symmetric).  Complex and
// hermitian files are rejected.
#include <cctype>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/hala_amd.h"

namespace hala_amd {
void set_error(const std::string &msg);
}

namespace {

struct mm_header {
    int array = 0;     // 0 coordinate, 1 array
    int field = 0;     // HALA_MM_REAL / INTEGER / PATTERN
    int symmetry = 0;  // HALA_MM_GENERAL / SYMMETRIC / SKEW
    int nrows = 0, ncols = 0;
    int64_t stored = 0;  // entries in the file
};

int fail(int code, const std::string &msg) {
    hala_amd::set_error(msg);
    return code;
}

std::string lower(std::string s) {
    for (char &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

// Reads the banner, comments and size line; leaves f at the first entry.
int read_header(FILE *f, const char *path, mm_header &h) {
    char line[4096];
    if (!std::fgets(line, sizeof line, f)) return fail(HALA_ERR_ARG, std::string("empty file: ") + path);
    char b[64] = {0}, obj[64] = {0}, fmt[64] = {0}, fld[64] = {0}, sym[64] = {0};
    if (std::sscanf(line, "%63s %63s %63s %63s %63s", b, obj, fmt, fld, sym) != 5 ||
        std::string(b) != "%%MatrixMarket")
        return fail(HALA_ERR_ARG, std::string("not a Matrix Market file: ") + path);
    const std::string o = lower(obj), fm = lower(fmt), fi = lower(fld), sy = lower(sym);
    if (o != "matrix") return fail(HALA_ERR_ARG, "Matrix Market object '" + o + "' is not a matrix");
    if (fm == "coordinate") h.array = 0;
    else if (fm == "array") h.array = 1;
    else return fail(HALA_ERR_ARG, "unknown Matrix Market format '" + fm + "'");
    if (fi == "real" || fi == "double") h.field = HALA_MM_REAL;
    else if (fi == "integer") h.field = HALA_MM_INTEGER;
    else if (fi == "pattern" && !h.array) h.field = HALA_MM_PATTERN;
    else return fail(HALA_ERR_ARG, "unsupported Matrix Market field '" + fi + "' (real, integer, pattern)");
    if (sy == "general") h.symmetry = HALA_MM_GENERAL;
    else if (sy == "symmetric") h.symmetry = HALA_MM_SYMMETRIC;
    else if (sy == "skew-symmetric") h.symmetry = HALA_MM_SKEW;
    else return fail(HALA_ERR_ARG, "unsupported Matrix Market symmetry '" + sy + "'");
    do {  // comments and blank lines
        if (!std::fgets(line, sizeof line, f)) return fail(HALA_ERR_ARG, "Matrix Market size line missing");
    } while (line[0] == '%' || line[strspn(line, " \t\r\n")] == '\0');
    long long r = 0, c = 0, z = 0;
    const int nread = h.array ? std::sscanf(line, "%lld %lld", &r, &c) : std::sscanf(line, "%lld %lld %lld", &r, &c, &z);
    if (nread != (h.array ? 2 : 3)) return fail(HALA_ERR_ARG, "bad Matrix Market size line");
    // the dimensions first: the array format's entry count is formed from them
    if (r < 0 || c < 0 || r > 0x7fffffff || c > 0x7fffffff) return fail(HALA_ERR_ARG, "Matrix Market sizes out of range");
    if (h.array) {
        if (h.symmetry == HALA_MM_GENERAL) z = r * c;
        else z = (h.symmetry == HALA_MM_SYMMETRIC) ? r * (r + 1) / 2 : (r > 0 ? r * (r - 1) / 2 : 0);
    }
    if (z < 0 || z >= 0x7fffffff) return fail(HALA_ERR_ARG, "Matrix Market sizes out of range");
    if (h.symmetry != HALA_MM_GENERAL && r != c) return fail(HALA_ERR_ARG, "symmetric Matrix Market file not square");
    h.nrows = (int)r;
    h.ncols = (int)c;
    h.stored = z;
    return HALA_OK;
}

struct file_closer {
    FILE *f;
    ~file_closer() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

int hala_mm_info(const char *path, int *nrows, int *ncols, int64_t *stored, int64_t *expanded, int *field,
                 int *symmetry) {
    if (!path) return fail(HALA_ERR_ARG, "hala_mm_info: null path");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(HALA_ERR_ARG, std::string("cannot open ") + path);
    file_closer fc{f};
    mm_header h;
    int st = read_header(f, path, h);
    if (st != HALA_OK) return st;
    if (nrows) *nrows = h.nrows;
    if (ncols) *ncols = h.ncols;
    if (stored) *stored = h.stored;
    if (field) *field = h.field;
    if (symmetry) *symmetry = h.symmetry;
    if (expanded) {
        // an upper bound without reading the entries: the off-diagonal ones are
        // counted exactly by hala_mm_read_coo
        *expanded = h.symmetry == HALA_MM_GENERAL ? h.stored : 2 * h.stored;
    }
    return HALA_OK;
}

int hala_mm_read_coo(const char *path, int expand, int64_t capacity, int *rows, int *cols, double *vals,
                     int64_t *count) {
    if (!path || !rows || !cols || !vals || !count) return fail(HALA_ERR_ARG, "hala_mm_read_coo: null argument");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(HALA_ERR_ARG, std::string("cannot open ") + path);
    file_closer fc{f};
    mm_header h;
    int st = read_header(f, path, h);
    if (st != HALA_OK) return st;
    int64_t q = 0;
    auto put = [&](int64_t i, int64_t j, double v) -> bool {
        if (q >= capacity) return false;
        rows[q] = (int)i;
        cols[q] = (int)j;
        vals[q] = v;
        q++;
        return true;
    };
    char line[4096];
    int64_t k = 0;
    // array format: column-major; symmetric ones hold the lower triangle (skew: strictly lower)
    int64_t ai = 0, aj = 0;
    if (h.array && h.symmetry == HALA_MM_SKEW) ai = 1;
    while (k < h.stored && std::fgets(line, sizeof line, f)) {
        if (line[0] == '%' || line[strspn(line, " \t\r\n")] == '\0') continue;
        long long i = 0, j = 0;
        double v = 1.0;
        char *p = line, *e = nullptr;
        if (h.array) {
            i = ai;
            j = aj;
            v = std::strtod(p, &e);
            if (e == p) return fail(HALA_ERR_ARG, "bad Matrix Market entry " + std::to_string(k + 1));
            if (++ai >= h.nrows) {
                aj++;
                ai = h.symmetry == HALA_MM_GENERAL ? 0 : (h.symmetry == HALA_MM_SYMMETRIC ? aj : aj + 1);
            }
        } else {
            // 1-based indices, range-checked before the shift (strtoll saturates)
            i = std::strtoll(p, &e, 10);
            if (e == p) return fail(HALA_ERR_ARG, "bad Matrix Market entry " + std::to_string(k + 1));
            p = e;
            j = std::strtoll(p, &e, 10);
            if (e == p) return fail(HALA_ERR_ARG, "bad Matrix Market entry " + std::to_string(k + 1));
            p = e;
            if (h.field != HALA_MM_PATTERN) {
                v = std::strtod(p, &e);
                if (e == p) return fail(HALA_ERR_ARG, "bad Matrix Market entry " + std::to_string(k + 1));
            }
            if (i < 1 || j < 1 || i > h.nrows || j > h.ncols)
                return fail(HALA_ERR_ARG, "Matrix Market entry " + std::to_string(k + 1) + " out of range");
            i -= 1;
            j -= 1;
        }
        k++;
        if (!put(i, j, v)) return fail(HALA_ERR_ARG, "hala_mm_read_coo: capacity too small");
        if (expand && h.symmetry != HALA_MM_GENERAL && i != j)
            if (!put(j, i, h.symmetry == HALA_MM_SKEW ? -v : v))
                return fail(HALA_ERR_ARG, "hala_mm_read_coo: capacity too small");
    }
    if (k != h.stored) return fail(HALA_ERR_ARG, "Matrix Market file ends after " + std::to_string(k) + " entries");
    *count = q;
    return HALA_OK;
}

int hala_mm_write_csr(const char *path, int nrows, int ncols, const int *pntr, const int *indx, const double *vals) {
    if (!path || !pntr || nrows < 0 || ncols < 0 || (pntr[nrows] > 0 && (!indx || !vals)))
        return fail(HALA_ERR_ARG, "hala_mm_write_csr: bad argument");
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(HALA_ERR_ARG, std::string("cannot create ") + path);
    file_closer fc{f};
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%d %d %d\n", nrows, ncols, pntr[nrows]);
    for (int i = 0; i < nrows; i++)
        for (int e = pntr[i]; e < pntr[i + 1]; e++) std::fprintf(f, "%d %d %.17g\n", i + 1, indx[e] + 1, vals[e]);
    if (std::ferror(f)) return fail(HALA_ERR_ARG, std::string("write failed: ") + path);
    return HALA_OK;
}

}  // extern "C"
