// stencils.cpp -- deterministic synthetic problems (host code).
//
// Random numbers come from splitmix64 (Steele, Lea, Flood 2014); a uniform
// double in [0,1) is (u >> 11) * 2^-53.  All matrices are sorted CSR with
// out-of-domain neighbours dropped (Dirichlet rows trimmed).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/hala_amd.h"

namespace {

struct splitmix64 {
    uint64_t s;
    explicit splitmix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
    double uniform(double lo, double hi) { return lo + (hi - lo) * u01(); }
    // Box-Muller, one normal per call (the sine branch is discarded)
    double normal() {
        double u1 = 1.0 - u01();  // (0, 1]
        double u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// 5-point operator on an nx x ny grid, row r = i + nx*j, columns sorted.
void laplace2d_fill(int nx, int ny, double dg, double off, int *pntr, int *indx, double *vals, int row0 = 0,
                    int64_t nnz0 = 0, bool absolute_pntr = false) {
    int64_t q = 0;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int r = i + nx * j;
            pntr[r] = (int)((absolute_pntr ? nnz0 : 0) + q);
            auto put = [&](int c, double v) {
                indx[q] = c;
                vals[q] = v;
                q++;
            };
            if (j > 0) put(r - nx, off);
            if (i > 0) put(r - 1, off);
            put(r, dg);
            if (i < nx - 1) put(r + 1, off);
            if (j < ny - 1) put(r + nx, off);
        }
    (void)row0;
    pntr[nx * ny] = (int)((absolute_pntr ? nnz0 : 0) + q);
}

inline int64_t laplace2d_nnz(int nx, int ny) { return 5ll * nx * ny - 2ll * nx - 2ll * ny; }

}  // namespace

extern "C" {

int hala_gen_laplace2d(int nx, int ny, double diag, double off, int *n, int64_t *nnz, int *pntr, int *indx,
                       double *vals) {
    if (nx < 1 || ny < 1) return HALA_ERR_ARG;
    if (n) *n = nx * ny;
    if (nnz) *nnz = laplace2d_nnz(nx, ny);
    if (pntr && indx && vals) laplace2d_fill(nx, ny, diag, off, pntr, indx, vals);
    return HALA_OK;
}

int hala_gen_laplace3d(int nx, int ny, int nz, int *n, int64_t *nnz, int *pntr, int *indx, double *vals) {
    if (nx < 1 || ny < 1 || nz < 1) return HALA_ERR_ARG;
    int64_t nn = (int64_t)nx * ny * nz;
    if (nn > 0x7fffffff) return HALA_ERR_ARG;
    int64_t z = 7 * nn - 2ll * ny * nz - 2ll * nx * nz - 2ll * nx * ny;
    if (z > 0x7fffffff) return HALA_ERR_ARG;
    if (n) *n = (int)nn;
    if (nnz) *nnz = z;
    if (!(pntr && indx && vals)) return HALA_OK;
    const int sxy = nx * ny;
    int64_t q = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int r = i + nx * j + sxy * k;
                pntr[r] = (int)q;
                auto put = [&](int c, double v) {
                    indx[q] = c;
                    vals[q] = v;
                    q++;
                };
                if (k > 0) put(r - sxy, -1.0);
                if (j > 0) put(r - nx, -1.0);
                if (i > 0) put(r - 1, -1.0);
                put(r, 6.0);
                if (i < nx - 1) put(r + 1, -1.0);
                if (j < ny - 1) put(r + nx, -1.0);
                if (k < nz - 1) put(r + sxy, -1.0);
            }
    pntr[nn] = (int)q;
    return HALA_OK;
}

// The 2D Poisson problem of HALA's GPU example (examples/example_gpu.cpp:407-478):
// the 5-point [4; -1] operator divided by dx^2, dx = 1/(num_side+1), and the
// forcing of the manufactured solution (x - x^4)(y - y^3) evaluated with the
// same expression so the table of examples/example_gpu.cpp:90-98 is reproducible.
int hala_gen_discretize_pde(int num_side, int *n, int64_t *nnz, int *pntr, int *indx, double *vals, double *f) {
    if (num_side < 2) return HALA_ERR_ARG;
    int ns = num_side;
    if (n) *n = ns * ns;
    if (nnz) *nnz = laplace2d_nnz(ns, ns);
    double dx = 1.0 / static_cast<double>(ns + 1);
    if (pntr && indx && vals) {
        laplace2d_fill(ns, ns, 4.0, -1.0, pntr, indx, vals);
        double h2 = dx * dx;
        for (int64_t q = 0; q < laplace2d_nnz(ns, ns); q++) vals[q] /= h2;
    }
    if (f) {
        for (int i = 0; i < ns; i++) {
            double y = dx * (i + 1);
            for (int j = 0; j < ns; j++) {
                double x = dx * (j + 1);
                f[i * ns + j] = 12.0 * x * x * (y - y * y * y) + 6 * y * (x - x * x * x * x);
            }
        }
    }
    return HALA_OK;
}

// 27-point anisotropic convection-diffusion on an m^3 grid.
//   D_c  = 10^(-2 + 4u): log-uniform in [1e-2, 1e2] per cell;
//   diffusion weight to neighbour offset o (26 of them):
//        g(o) = (|ox| ax + |oy| ay + |oz| az) / |o|_1^2,  (ax, ay, az) = (1, 0.3, 0.1),
//        times the harmonic mean of D over the two cells (D_c outside the domain);
//   convection V_c = peclet * D_c * v, v ~ U[-1,1]^3: first-order upwind per axis
//        (the upstream face neighbour gets -|V_a|, the diagonal +|V_a|);
//   diagonal = sum of all 26 diffusion weights (Dirichlet neighbours included)
//        + sum |V_a|, so every row is weakly diagonally dominant (an M-matrix).
int hala_gen_convdiff27(int m, uint64_t seed, double peclet, int *n, int64_t *nnz, int *pntr, int *indx,
                        double *vals) {
    if (m < 2) return HALA_ERR_ARG;
    int64_t nn = (int64_t)m * m * m;
    int64_t t = 3ll * m - 2;
    int64_t z = t * t * t;
    if (nn > 0x7fffffff || z > 0x7fffffff) return HALA_ERR_ARG;
    if (n) *n = (int)nn;
    if (nnz) *nnz = z;
    if (!(pntr && indx && vals)) return HALA_OK;
    splitmix64 g(seed);
    std::vector<double> D(nn), V(3 * nn);
    for (int64_t c = 0; c < nn; c++) {
        D[c] = std::pow(10.0, -2.0 + 4.0 * g.u01());
        for (int a = 0; a < 3; a++) V[3 * c + a] = peclet * D[c] * g.uniform(-1.0, 1.0);
    }
    const double ax[3] = {1.0, 0.3, 0.1};
    const int64_t sxy = (int64_t)m * m;
    int64_t q = 0;
    for (int k = 0; k < m; k++)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) {
                int64_t r = i + (int64_t)m * j + sxy * k;
                pntr[r] = (int)q;
                double dc = D[r];
                double diag = 0.0;
                int64_t qdiag = -1;
                for (int oz = -1; oz <= 1; oz++)
                    for (int oy = -1; oy <= 1; oy++)
                        for (int ox = -1; ox <= 1; ox++) {
                            int ii = i + ox, jj = j + oy, kk = k + oz;
                            bool inside = ii >= 0 && ii < m && jj >= 0 && jj < m && kk >= 0 && kk < m;
                            if (ox == 0 && oy == 0 && oz == 0) {
                                qdiag = q;
                                indx[q] = (int)r;
                                vals[q] = 0.0;
                                q++;
                                continue;
                            }
                            int l1 = std::abs(ox) + std::abs(oy) + std::abs(oz);
                            double gw = (std::abs(ox) * ax[0] + std::abs(oy) * ax[1] + std::abs(oz) * ax[2]) /
                                        (double)(l1 * l1);
                            int64_t nb = inside ? ii + (int64_t)m * jj + sxy * kk : -1;
                            double dn = inside ? D[nb] : dc;
                            double w = gw * (2.0 * dc * dn / (dc + dn));
                            diag += w;
                            if (!inside) continue;
                            double v = -w;
                            if (l1 == 1) {  // face neighbour: upwind convection along its axis
                                int a = ox ? 0 : (oy ? 1 : 2);
                                int s = ox + oy + oz;  // +1 or -1
                                double va = V[3 * r + a];
                                // upstream neighbour is at -sign(va) along the axis
                                if ((va > 0.0 && s < 0) || (va < 0.0 && s > 0)) v -= std::fabs(va);
                            }
                            indx[q] = (int)nb;
                            vals[q] = v;
                            q++;
                        }
                for (int a = 0; a < 3; a++) diag += std::fabs(V[3 * r + a]);
                vals[qdiag] = diag;
            }
    pntr[nn] = (int)q;
    return HALA_OK;
}

int hala_gen_uniform(uint64_t seed, int64_t n, double lo, double hi, double *out) {
    if (n < 0 || !out) return HALA_ERR_ARG;
    splitmix64 g(seed);
    for (int64_t i = 0; i < n; i++) out[i] = g.uniform(lo, hi);
    return HALA_OK;
}

// nsys independent systems: sides nx, ny uniform in [side_lo, side_hi], a 5-point
// operator [4 + s; -1] with shift s ~ U[0.1, 1.0] per system.  Packed layout:
// row_off[nsys+1] and nnz_off[nsys+1] are the system offsets, pntr[n_total+1]
// holds absolute positions in indx/vals, indx holds columns local to the system,
// dvals[n_total] the diagonal values (the Jacobi preconditioner).
int hala_gen_batch_laplace(int nsys, uint64_t seed, int side_lo, int side_hi, int *n_total, int64_t *nnz_total,
                           int *sys_rows, int *row_off, int64_t *nnz_off, int *pntr, int *indx, double *vals,
                           double *dvals) {
    if (nsys < 1 || side_lo < 2 || side_hi < side_lo) return HALA_ERR_ARG;
    splitmix64 g(seed);
    std::vector<int> nxs(nsys), nys(nsys);
    std::vector<double> sh(nsys);
    int64_t nt = 0, zt = 0;
    for (int s = 0; s < nsys; s++) {
        int span = side_hi - side_lo + 1;
        nxs[s] = side_lo + (int)(g.u01() * span);
        nys[s] = side_lo + (int)(g.u01() * span);
        sh[s] = g.uniform(0.1, 1.0);
        nt += (int64_t)nxs[s] * nys[s];
        zt += laplace2d_nnz(nxs[s], nys[s]);
    }
    if (nt > 0x7fffffff || zt > 0x7fffffff) return HALA_ERR_ARG;
    if (n_total) *n_total = (int)nt;
    if (nnz_total) *nnz_total = zt;
    if (!(sys_rows && row_off && nnz_off && pntr && indx && vals && dvals)) return HALA_OK;
    int64_t r0 = 0, z0 = 0;
    row_off[0] = 0;
    nnz_off[0] = 0;
    for (int s = 0; s < nsys; s++) {
        int ns = nxs[s] * nys[s];
        sys_rows[s] = ns;
        laplace2d_fill(nxs[s], nys[s], 4.0 + sh[s], -1.0, pntr + r0, indx + z0, vals + z0, 0, z0, true);
        for (int i = 0; i < ns; i++) dvals[r0 + i] = 4.0 + sh[s];
        r0 += ns;
        z0 += laplace2d_nnz(nxs[s], nys[s]);
        row_off[s + 1] = (int)r0;
        nnz_off[s + 1] = z0;
    }
    pntr[nt] = (int)zt;
    return HALA_OK;
}

// 7-point variable-coefficient diffusion on an m^3 grid with one conductivity
// kappa_f = exp(z), z ~ N(0,1), per cell face (boundary faces included).
// Faces: x-faces i + (m+1)(j + m k), i in [0,m]; then y-faces i + m(j + (m+1)k);
// then z-faces i + m(j + m k); ncoef = 3 (m+1) m^2.  Off-diagonal (p,q) = -kappa
// of the shared face; diagonal = sum of kappa over the six faces of the cell.
// Coefficient map (for the chain rule onto kappa): for face c the entries
// cptr[c]..cptr[c+1] list (cnnz, ccoef) = the nnz positions it feeds and +-1.
int hala_gen_diffusion3d(int m, uint64_t seed, int *n, int64_t *nnz, int *ncoef, int64_t *nmap, double *kappa,
                         int *pntr, int *indx, double *vals, int *cptr, int *cnnz, double *ccoef) {
    if (m < 2) return HALA_ERR_ARG;
    int64_t nn = (int64_t)m * m * m;
    int64_t fx = (int64_t)(m + 1) * m * m;
    int64_t nf = 3 * fx;
    int64_t z = 7 * nn - 6ll * m * m;
    int64_t interior = nf - 6ll * m * m;
    int64_t nm = 4 * interior + (nf - interior);
    if (n) *n = (int)nn;
    if (nnz) *nnz = z;
    if (ncoef) *ncoef = (int)nf;
    if (nmap) *nmap = nm;
    if (!(kappa && pntr && indx && vals && cptr && cnnz && ccoef)) return HALA_OK;
    splitmix64 g(seed);
    for (int64_t f = 0; f < nf; f++) kappa[f] = std::exp(g.normal());
    const int64_t sxy = (int64_t)m * m;
    auto face = [&](int a, int i, int j, int k) -> int64_t {
        if (a == 0) return i + (int64_t)(m + 1) * (j + (int64_t)m * k);
        if (a == 1) return fx + i + (int64_t)m * (j + (int64_t)(m + 1) * k);
        return 2 * fx + i + (int64_t)m * (j + (int64_t)m * k);
    };
    // pattern and values; remember each face's entries
    std::vector<std::vector<std::pair<int64_t, double>>> fmap(nf);
    int64_t q = 0;
    for (int k = 0; k < m; k++)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) {
                int64_t r = i + (int64_t)m * j + sxy * k;
                pntr[r] = (int)q;
                // faces of the cell: (axis, lower/upper face id)
                int64_t fl[3] = {face(0, i, j, k), face(1, i, j, k), face(2, i, j, k)};
                int64_t fu[3] = {face(0, i + 1, j, k), face(1, i, j + 1, k), face(2, i, j, k + 1)};
                auto put = [&](int64_t c, int64_t f) {
                    indx[q] = (int)c;
                    vals[q] = -kappa[f];
                    fmap[f].push_back({q, -1.0});
                    q++;
                };
                if (k > 0) put(r - sxy, fl[2]);
                if (j > 0) put(r - m, fl[1]);
                if (i > 0) put(r - 1, fl[0]);
                int64_t qd = q;
                indx[q] = (int)r;
                q++;
                double d = 0.0;
                for (int a = 0; a < 3; a++) {
                    d += kappa[fl[a]];
                    fmap[fl[a]].push_back({qd, 1.0});
                    d += kappa[fu[a]];
                    fmap[fu[a]].push_back({qd, 1.0});
                }
                vals[qd] = d;
                if (i < m - 1) put(r + 1, fu[0]);
                if (j < m - 1) put(r + m, fu[1]);
                if (k < m - 1) put(r + sxy, fu[2]);
            }
    pntr[nn] = (int)q;
    int64_t e = 0;
    for (int64_t f = 0; f < nf; f++) {
        cptr[f] = (int)e;
        for (auto &pr : fmap[f]) {
            cnnz[e] = (int)pr.first;
            ccoef[e] = pr.second;
            e++;
        }
    }
    cptr[nf] = (int)e;
    return HALA_OK;
}

// Gaussian source exp(-|x - c|^2 / (2 sigma^2)) at cell centres (i+1/2)/m, c = (1/2,1/2,1/2).
int hala_gen_gaussian_source(int m, double sigma, double *f) {
    if (m < 1 || !f) return HALA_ERR_ARG;
    const int64_t sxy = (int64_t)m * m;
    for (int k = 0; k < m; k++)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) {
                double x = (i + 0.5) / m - 0.5, y = (j + 0.5) / m - 0.5, zz = (k + 0.5) / m - 0.5;
                f[i + (int64_t)m * j + sxy * k] = std::exp(-(x * x + y * y + zz * zz) / (2.0 * sigma * sigma));
            }
    return HALA_OK;
}

}  // extern "C"
