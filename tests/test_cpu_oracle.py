"""CPU tests of the oracle (no GPU): pins against the reference's own test
vectors and published table, against HALA's compiled kernels (fixtures and,
where the reference tree is present, the live build), and against what the
mathematics fixes (closed forms, invariants, brute force on tiny inputs)."""
import json
import os

import numpy as np
import pytest

import oracle
from hala_amd import stencils as st
from tests.conftest import bits

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dense(p, i, v, nrows, ncols=None):
    A = np.zeros((nrows, ncols or nrows))
    for r in range(nrows):
        A[r, i[p[r]:p[r + 1]]] = v[p[r]:p[r + 1]]
    return A


TRI5 = (np.array([0, 2, 5, 8, 11, 13], np.int32), np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4], np.int32),
        np.array([2, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 2], np.float64))


# ---------------------------------------------------------------- published and test-suite vectors
def test_published_pcg_table():
    table = json.load(open(os.path.join(GOLDEN, "published_pcg_table.json")))
    for row in table["rows"]:
        p, i, v, f = st.discretize_pde(row["num_side"])
        dg, bad = oracle.diag(p, i)
        assert bad == -1
        it, u, _ = oracle.cg(p, i, v, f, tol=1e-8, prec=oracle.PREC_ILU, mpntr=p, mindx=i, mdiag=dg,
                             mdata=oracle.ilu0(p, i, v, dg))
        assert it == row["iterations"]
        assert float("%.6e" % st.l2_error(u)) == row["l2_error"]


def test_sparse_gemv_reference_vectors():
    # tests/sparse_tests.hpp:165-191 (alpha = 2, N and T) with hala_rnd-style integer x
    p, i, v = TRI5
    x = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    A = dense(p, i, v, 5)
    assert np.array_equal(oracle.spmv("N", 5, 5, 2.0, p, i, v, x, 0.0, np.zeros(5)), 2.0 * A @ x)
    assert np.array_equal(oracle.spmv("T", 5, 5, 2.0, p, i, v, x, 0.0, np.zeros(5)), 2.0 * A.T @ x)
    # rectangular 5x6 of tests/sparse_tests.hpp:138-146
    p = np.array([0, 3, 7, 9, 13, 16], np.int32)
    i = np.array([0, 2, 5, 0, 1, 2, 5, 2, 3, 0, 2, 3, 5, 1, 3, 4], np.int32)
    v = np.array([2.0, 4.0, 5.0, 1.0, 3.0, 1.0, 5.0, 4.0, 1.0, 3.0, 4.0, 4.0, 1.0, 5.0, 1.0, 3.0])
    A = dense(p, i, v, 5, 6)
    x = np.arange(6.0)
    assert np.array_equal(oracle.spmv("N", 5, 6, 2.0, p, i, v, x, 0.0, np.zeros(5)), 2.0 * A @ x)
    # cmake/post_install_test.sh: diagonal 3x3 -> {4, 10, 18}
    y = oracle.spmv("N", 3, 3, 1.0, [0, 1, 2, 3], [0, 1, 2], [4.0, 5.0, 6.0], [1.0, 2.0, 3.0], 0.0, np.zeros(3))
    assert list(y) == [4.0, 10.0, 18.0]


def test_split_diag_reference():
    # tests/sparse_tests.hpp:368-386
    p = np.array([0, 2, 5, 8, 11, 13, 17], np.int32)
    i = np.array([0, 3, 0, 1, 4, 2, 3, 4, 0, 3, 4, 2, 4, 0, 1, 3, 5], np.int32)
    d, bad = oracle.diag(p, i)
    assert list(d) == [0, 3, 5, 9, 12, 16] and bad == -1


def test_ilu_reference_known_answer():
    # tests/sparse_tests.hpp:423-440
    p = np.array([0, 3, 6, 9], np.int32)
    i = np.array([0, 1, 2, 0, 1, 2, 0, 1, 2], np.int32)
    v = np.array([3.0, 2.0, 1.0, 2.0, 4.0, 3.0, 2.0, 1.0, 6.0])
    d, _ = oracle.diag(p, i)
    assert list(d) == [0, 4, 8]
    f = oracle.ilu0(p, i, v, d)
    ref = np.array([3.0, 2.0, 1.0, 2.0 / 3.0, 8.0 / 3.0, 7.0 / 3.0, 2.0 / 3.0, -0.125, 5.625])
    assert np.abs(f - ref).sum() < 1e-14
    x = oracle.ilu_apply(p, i, d, f, [1.0, 2.0, 3.0])
    assert np.abs(x - [1 / 9, 1 / 9, 4 / 9]).sum() < 1e-14


def test_trsv_reference_cases():
    # tests/sparse_tests.hpp:259-311: alpha = 0.5, B = 2 op(A) xref, xref from hala_rnd = -4
    A = np.array([1.0, 4.0, 5.0, 0.0, 2.0, 0.0, 0.0, 0.0, 3.0]).reshape(3, 3, order="F")
    xref = np.array([-3.0, -2.0, -1.0])
    L = ([0, 1, 3, 5], [0, 0, 1, 0, 2], [1.0, 4.0, 2.0, 5.0, 3.0])
    U = ([0, 3, 4, 5], [0, 1, 2, 1, 2], [1.0, 4.0, 5.0, 2.0, 3.0])
    for uplo, trans, M, Aop in [("L", "N", L, A), ("L", "T", L, A.T), ("U", "N", U, A.T), ("U", "T", U, A)]:
        x = oracle.trsv(uplo, trans, "N", 0.5, *M, 2.0 * Aop @ xref)
        assert np.abs(x - xref).sum() < 1e-14, (uplo, trans)
    Au = np.array([1.0, 1.0, 2.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]).reshape(3, 3, order="F")
    x = oracle.trsv("L", "N", "U", 0.5, [0, 1, 3, 5], [0, 0, 1, 0, 2], [2.0, 1.0, 3.0, 2.0, 4.0], 2.0 * Au @ xref)
    assert np.abs(x - xref).sum() < 1e-14


def test_cg_reference_cases():
    # tests/solvers_tests.hpp:12-55
    p, i, v = TRI5
    xref = np.arange(1.0, 6.0)
    b = dense(p, i, v, 5) @ xref
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    it, x, _ = oracle.cg(p, i, v, b, max_iter=2, prec=oracle.PREC_ILU, mpntr=p, mindx=i, mdiag=d, mdata=f)
    assert np.abs(x - xref).sum() < 1e-14
    p = np.array([0, 3, 6, 9, 12, 15], np.int32)
    i = np.array([0, 1, 4, 0, 1, 2, 1, 2, 3, 2, 3, 4, 0, 3, 4], np.int32)
    v = np.array([2, 1, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 1, 2], np.float64)
    b = dense(p, i, v, 5) @ xref
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    it, x, _ = oracle.cg(p, i, v, b, x0=[1.0, 3.0, -4.0, -7.0, 11.0], tol=1e-9, prec=oracle.PREC_ILU, mpntr=p,
                         mindx=i, mdiag=d, mdata=f)
    assert np.abs(x - xref).sum() < 1e-9 * np.linalg.norm(xref) * 10


def test_gmres_reference_cases():
    # tests/solvers_tests.hpp:105-137
    p, i, v = TRI5
    xref = np.arange(1.0, 6.0)
    b = dense(p, i, v, 5) @ xref
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    _, x, _ = oracle.gmres(p, i, v, b, tol=0.0, restart=2, max_outer=2, dg=d, ilu=f)
    assert np.abs(x - xref).sum() < 1e-14
    p = np.array([0, 2, 5, 8, 11, 12], np.int32)
    i = np.array([0, 1, 0, 1, 3, 1, 2, 3, 0, 2, 3, 4], np.int32)
    v = np.array([2.0, 1.0, 1.0, 3.0, 1.0, 1.0, 4.0, 1.0, 1.0, 1.0, 5.0, 2.0])
    b = dense(p, i, v, 5) @ xref
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    _, x, _ = oracle.gmres(p, i, v, b, tol=1e-9, restart=6, dg=d, ilu=f)
    assert np.abs(x - xref).sum() < 1e-8


NONSYM5 = (np.array([0, 2, 5, 8, 11, 12], np.int32), np.array([0, 1, 0, 1, 3, 1, 2, 3, 0, 2, 3, 4], np.int32),
           np.array([2.0, 1.0, 1.0, 3.0, 1.0, 1.0, 4.0, 1.0, 1.0, 1.0, 5.0, 2.0]))


def _drotg_r(a, b):
    """r of the reference-BLAS drotg: +-sqrt(a^2 + b^2), signed like the larger of |a|, |b|."""
    roe = a if abs(a) > abs(b) else b
    return np.copysign(np.hypot(a, b), roe if roe != 0 else 1.0)


def test_gmres1_cycle_closed_form_unrotated_last_z():
    # One GMRES(1) cycle from x0 = 0 (identity preconditioner).  With w = b/|b|,
    # h = <w, A w> and e = |A w - h w|, the Givens rotation of (h, e) gives
    # rho = drotg's r and c = h / rho.  The least-squares step is |b| c / rho;
    # HALA never applies the cycle's last rotation to Z
    # (hex/solvers/hala_solvers_gmres.hpp:210-216), so its step is |b| / rho.
    p, i, v = NONSYM5
    A = dense(p, i, v, 5)
    b = np.array([1.0, -2.0, 0.5, 3.0, -1.0])
    _, x, _ = oracle.gmres(p, i, v, b, tol=0.0, restart=1, max_outer=1, prec=oracle.PREC_IDENTITY)
    beta = np.linalg.norm(b)
    w = b / beta
    h = w @ (A @ w)
    e = np.linalg.norm(A @ w - h * w)
    rho = _drotg_r(h, e)
    c = h / rho
    assert abs(c) < 0.95  # the two variants are far apart on this system
    np.testing.assert_allclose(x, (beta / rho) * w, rtol=1e-13, atol=1e-15)
    assert np.linalg.norm(x - (beta * c / rho) * w) > 1e-2 * np.linalg.norm(x)


@pytest.mark.parametrize("k", [2, 3, 4])
def test_gmres_cycle_differs_from_least_squares_by_last_rotation(k):
    # One GMRES(k) cycle of k steps from x0 = 0.  On the Krylov space W_k the
    # least-squares step y* solves R y = g (A W_k = Q R, g = Q^T b); HALA keeps
    # Z's last entry unrotated, i.e. g_k / c_k instead of g_k, so
    # x - x* = (1/c_k - 1) W_k R^-1 g_k e_k, with |c_k| = sqrt(R_kk^2 - e_k^2) / |R_kk|
    # and e_k the norm of A w_k outside W_k.  The rotated variant would give x = x*.
    p, i, v = st.convdiff27(3, seed=2)
    n = len(p) - 1
    A = dense(p, i, v, n)
    b = st.uniform(11, n)
    _, x, _ = oracle.gmres(p, i, v, b, tol=0.0, restart=k, max_outer=1, prec=oracle.PREC_IDENTITY)
    K = np.empty((n, k))
    K[:, 0] = b
    for j in range(1, k):
        K[:, j] = A @ K[:, j - 1]
    W, _ = np.linalg.qr(K)
    Q, R = np.linalg.qr(A @ W)
    g = Q.T @ b
    xs = W @ np.linalg.solve(R, g)
    d = W @ np.linalg.solve(R, np.eye(k)[:, -1] * g[-1])
    wk = W[:, -1]
    aw = A @ wk
    e_k = np.linalg.norm(aw - W @ (W.T @ aw))
    c_abs = np.sqrt(R[-1, -1] ** 2 - e_k**2) / abs(R[-1, -1])
    t = (x - xs) @ d / (d @ d)
    np.testing.assert_allclose(x - xs, t * d, atol=1e-9 * np.linalg.norm(x))  # along the predicted direction
    assert min(abs(t - (1 / c_abs - 1)), abs(t - (-1 / c_abs - 1))) < 1e-7 * max(1.0, abs(t))
    assert abs(t) > 1e-3  # and not the least-squares step


# ---------------------------------------------------------------- HALA's compiled kernels
REF_CASES = {
    "lap2d_16x12": lambda: st.laplace2d(16, 12),
    "lap3d_7": lambda: st.laplace3d(7, 7, 7),
    "cd27_6": lambda: st.convdiff27(6, seed=3),
    "pde_10": lambda: st.discretize_pde(10)[:3],
    "diff3d_5": lambda: (lambda d: (d["pntr"], d["indx"], d["vals"]))(st.diffusion3d(5, seed=5)),
}


@pytest.mark.parametrize("name", sorted(REF_CASES))
def test_oracle_matches_hala_fixtures(name):
    g = np.load(os.path.join(GOLDEN, "hala_ref_kernels.npz"))
    p, i, v = REF_CASES[name]()
    n = len(p) - 1
    d, _ = oracle.diag(p, i)
    assert np.array_equal(d, g[name + "/diag"])
    f = oracle.ilu0(p, i, v, d)
    assert np.array_equal(bits(f), bits(g[name + "/ilu"]))
    x = st.uniform(1234, n)
    assert np.array_equal(bits(oracle.ilu_apply(p, i, d, f, x)), bits(g[name + "/apply"]))
    y = st.uniform(99, n)
    assert np.array_equal(bits(oracle.spmv("N", n, n, 1.5, p, i, v, x, -0.75, y)), bits(g[name + "/spmv"]))
    x32 = x.astype(np.float32)
    assert np.array_equal(bits(oracle.ilu_apply_f32(p, i, d, f.astype(np.float32), x32)), bits(g[name + "/apply32"]))
    assert np.array_equal(bits(oracle.spmv_f32(p, i, v.astype(np.float32), x32)), bits(g[name + "/spmv32"]))


def test_oracle_matches_live_reference_build():
    # resolved inside the test: collecting this module loads nothing built from the reference
    R, P = oracle.ref(), oracle.P
    if R is None:
        pytest.skip("reference tree not available to build oracle/_ref")
    for p, i, v in (st.laplace3d(14, 12, 10), st.convdiff27(10, seed=11), st.discretize_pde(33)[:3]):
        n = len(p) - 1
        d, _ = oracle.diag(p, i)
        dr = np.empty(n, np.int32)
        R.hala_ref_diag(n, P(p), P(i), P(dr))
        assert np.array_equal(d, dr)
        fr = v.copy()
        R.hala_ref_ilu0(n, P(p), P(i), P(d), P(fr))
        assert np.array_equal(bits(oracle.ilu0(p, i, v, d)), bits(fr))


# ---------------------------------------------------------------- mathematics
def test_cdot_exact_on_integers_and_definition():
    rng = np.random.default_rng(0)
    for n in [0, 1, 255, 256, 257, 70000]:
        x = rng.integers(-1000, 1000, n).astype(np.float64)
        y = rng.integers(-1000, 1000, n).astype(np.float64)
        assert oracle.cdot(x, y) == float(np.dot(x.astype(np.int64), y.astype(np.int64)))
    # the canonical order differs from a sequential sum: it must match a direct restatement
    x = st.uniform(5, 1000)
    part = []
    for b0 in range(0, 1024, 256):
        a = np.zeros(256)
        seg = x[b0:b0 + 256] * x[b0:b0 + 256]
        a[: len(seg)] = seg
        off = 128
        while off >= 1:
            a[:off] = a[:off] + a[off:2 * off]
            off //= 2
        part.append(a[0])
    s = np.zeros(256)
    for t in range(256):
        acc = 0.0
        for k in range(t, len(part), 256):
            acc = acc + part[k]
        s[t] = acc
    off = 128
    while off >= 1:
        s[:off] = s[:off] + s[off:2 * off]
        off //= 2
    assert oracle.cdot(x, x) == s[0]


def test_ilu_defining_invariant_and_loop_orders():
    # (L U)_ij = A_ij on the pattern; KIJ (HALA's order) == IKJ bit for bit
    for p, i, v in (st.laplace2d(9, 7), st.convdiff27(4, seed=2), st.laplace3d(5, 4, 3)):
        n = len(p) - 1
        d, _ = oracle.diag(p, i)
        f = oracle.ilu0(p, i, v, d)
        assert np.array_equal(bits(f), bits(oracle.ilu0(p, i, v, d, kij=True)))
        F = dense(p, i, f, n)
        L = np.tril(F, -1) + np.eye(n)
        U = np.triu(F)
        LU = L @ U
        A = dense(p, i, v, n)
        mask = A != 0
        np.testing.assert_allclose(LU[mask], A[mask], rtol=1e-12, atol=1e-12 * np.abs(A).max())
        # apply == solve with the dense factors
        x = st.uniform(3, n)
        np.testing.assert_allclose(oracle.ilu_apply(p, i, d, f, x), np.linalg.solve(U, np.linalg.solve(L, x)),
                                   rtol=1e-10)
        np.testing.assert_allclose(oracle.ilu_apply_t(p, i, d, f, x), np.linalg.solve((L @ U).T, x), rtol=1e-10)


def test_tridiagonal_ilu_is_exact_lu():
    p, i, v = TRI5
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    F = dense(p, i, f, 5)
    L = np.tril(F, -1) + np.eye(5)
    U = np.triu(F)
    np.testing.assert_allclose(L @ U, dense(p, i, v, 5), atol=1e-15)


def _mirror_certificate(p, i):
    """(U's level recurrence holds for the candidate nlev_L - 1 - lev_L, candidate == U's levels)."""
    n = len(p) - 1
    dg, _ = oracle.diag(p, i)
    lo, _, _ = oracle.levels(p, i, dg, False)
    up, _, _ = oracle.levels(p, i, dg, True)
    cand = lo.max() - lo
    ok = all((0 if dg[r] + 1 == p[r + 1] else 1 + int(cand[i[dg[r] + 1:p[r + 1]]].max())) == cand[r]
             for r in range(n))
    return ok, bool(np.array_equal(up, cand))


def test_level_mirror_certificate():
    # U's levels are taken from L's (nlev_L - 1 - lev_L) when that candidate satisfies U's own
    # recurrence (0 without upper entries, else 1 + the largest over them); the recurrence has one
    # solution, so the check passes exactly when the candidate is U's level set.  It holds on the
    # stencils (every row lies on a longest chain) and fails on rows off it.
    for p, i, v in (st.laplace2d(9, 7), st.laplace3d(6, 5, 4), st.convdiff27(5, seed=1)):
        assert _mirror_certificate(p, i) == (True, True)
    # a 2-D grid beside a disconnected 3-row chain: the chain's rows are off the longest chain
    p, i, v = st.laplace2d(8, 8)
    n0 = len(p) - 1
    rows = [list(i[p[r]:p[r + 1]]) for r in range(n0)] + [[n0, n0 + 1], [n0, n0 + 1, n0 + 2], [n0 + 1, n0 + 2]]
    P = np.zeros(len(rows) + 1, np.int32)
    P[1:] = np.cumsum([len(c) for c in rows])
    I = np.concatenate([np.array(c, np.int32) for c in rows])
    assert _mirror_certificate(P, I) == (False, False)


def test_level_set_closed_forms():
    # natural x-fastest ordering: 5-pt level = i + j; 7-pt = i + j + k; 27-pt = i + 2j + 4k
    for (nx, ny), (p, i, v) in [((7, 5), st.laplace2d(7, 5))]:
        d, _ = oracle.diag(p, i)
        lev, lptr, perm = oracle.levels(p, i, d, False)
        jj, ii = np.divmod(np.arange(nx * ny), nx)
        assert np.array_equal(lev, ii + jj)
        assert len(lptr) - 1 == nx + ny - 1
    p, i, v = st.laplace3d(6, 5, 4)
    d, _ = oracle.diag(p, i)
    lev, lptr, perm = oracle.levels(p, i, d, False)
    r = np.arange(6 * 5 * 4)
    assert np.array_equal(lev, r % 6 + (r // 6) % 5 + r // 30)
    levu, _, _ = oracle.levels(p, i, d, True)
    assert np.array_equal(levu, (5 - r % 6) + (4 - (r // 6) % 5) + (3 - r // 30))
    m = 5
    p, i, v = st.convdiff27(m, seed=1)
    d, _ = oracle.diag(p, i)
    lev, lptr, perm = oracle.levels(p, i, d, False)
    r = np.arange(m ** 3)
    assert np.array_equal(lev, r % m + 2 * ((r // m) % m) + 4 * (r // (m * m)))
    # perm lists rows by (level, row) and level_ptr counts them
    assert np.array_equal(perm, np.lexsort((r, lev)))
    assert np.array_equal(np.diff(lptr), np.bincount(lev))


def test_rotg_tpsv_definitions():
    for a, b in [(3.0, 4.0), (-2.0, 1e-3), (0.0, 5.0), (1e-300, 1e-300), (0.0, 0.0)]:
        r, z, c, s = oracle.rotg(a, b)
        assert abs(c * c + s * s - 1.0) < 1e-15 or (a == 0 and b == 0)
        assert abs(-s * a + c * b) <= 1e-15 * max(abs(a), abs(b), 1e-300)
        assert abs(c * a + s * b - r) <= 1e-15 * max(abs(r), 1e-300)
    rng = np.random.default_rng(1)
    n = 7
    U = np.triu(rng.uniform(1, 2, (n, n)))
    ap = np.concatenate([U[: j + 1, j] for j in range(n)])
    x = rng.uniform(-1, 1, n)
    np.testing.assert_allclose(oracle.tpsv_un(ap, x), np.linalg.solve(U, x), rtol=1e-12)


def test_cg_and_gmres_converge_to_true_solution():
    p, i, v = st.laplace2d(20, 17)
    n = len(p) - 1
    b = st.uniform(6, n)
    A = dense(p, i, v, n)
    xs = np.linalg.solve(A, b)
    it, x, h = oracle.cg(p, i, v, b, tol=1e-12 * np.linalg.norm(b))
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10
    p, i, v = st.convdiff27(6, seed=4)
    n = len(p) - 1
    b = st.uniform(7, n)
    xs = np.linalg.solve(dense(p, i, v, n), b)
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    _, x, h = oracle.gmres(p, i, v, b, tol=1e-13, restart=10, dg=d, ilu=f)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-9
    assert np.all(np.diff(h[:5]) <= 0)  # Givens estimates do not increase within a cycle


@pytest.mark.parametrize("fp32", [True, False])
def test_refinement_reaches_fp64_accuracy(fp32):
    p, i, v = st.convdiff27(10, seed=3)
    n = len(p) - 1
    b = st.uniform(8, n)
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    tot, x, hist = oracle.gmres_ir(p, i, v, d, f, b, rtol=1e-10, fp32=fp32)
    assert hist[-1] < 1e-10 and len(hist) >= 2
    # every refinement step cuts the true residual by well over an order of magnitude
    assert np.all(hist[1:] < 0.05 * hist[:-1])
    xs = np.linalg.solve(dense(p, i, v, n), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


def _lu_dense(p, i, f, n):
    a = dense(p, i, f, n)
    return np.tril(a, -1), np.triu(a, 1), np.diag(a).copy()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", ["lap2d_12x9", "cd27_8"])
def test_jacobi_ilu_apply_pins(case, prec):
    """ILU applied by k Jacobi sweeps per triangle (FP32 or FP64 factor): equals the exact apply
    of the same precision once k reaches the level depth (same per-row term order); one sweep
    matches the dense algebra y = x - (L-I)x, z = (y - (U-D) y/d) / d; the error falls with
    every sweep."""
    p, i, v = st.laplace2d(12, 9) if case == "lap2d_12x9" else st.convdiff27(8, seed=3)
    n = len(p) - 1
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    if prec == "f32":
        fp = f.astype(np.float32)
        x = st.uniform(9, n).astype(np.float32)
        exact = oracle.ilu_apply_f32(p, i, d, fp, x)
        jac = lambda k: oracle.ilu_apply_f32_jacobi(p, i, d, fp, x, k)  # noqa: E731
        rtol = 1e-5
    else:
        fp = f
        x = st.uniform(9, n)
        exact = oracle.ilu_apply(p, i, d, f, x)
        jac = lambda k: oracle.ilu_apply_jacobi(p, i, d, f, x, k)  # noqa: E731
        rtol = 1e-13
    nl = max(len(oracle.levels(p, i, d, up)[1]) - 1 for up in (False, True))
    assert np.array_equal(bits(jac(nl)), bits(exact))
    lo, up, dg = _lu_dense(p, i, fp.astype(np.float64), n)
    x64 = x.astype(np.float64)
    y = x64 - lo @ x64
    z = (y - up @ (y / dg)) / dg
    one = jac(1).astype(np.float64)
    assert np.linalg.norm(one - z) <= rtol * np.linalg.norm(z)
    errs = [np.linalg.norm(jac(k).astype(np.float64) - exact.astype(np.float64)) for k in range(1, 7)]
    assert all(b < a for a, b in zip(errs, errs[1:])), errs


@pytest.mark.parametrize("fp32", [True, False])
def test_refinement_with_jacobi_inner_reaches_fp64_accuracy(fp32):
    p, i, v = st.convdiff27(10, seed=3)
    n = len(p) - 1
    b = st.uniform(8, n)
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    tot_exact, _, _ = oracle.gmres_ir(p, i, v, d, f, b, rtol=1e-10, fp32=fp32)
    tot, x, hist = oracle.gmres_ir(p, i, v, d, f, b, rtol=1e-10, fp32=fp32, jacobi=3)
    assert hist[-1] < 1e-10
    assert tot >= tot_exact  # an inexact preconditioner needs at least as many inner steps
    xs = np.linalg.solve(dense(p, i, v, n), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


def test_batch_cg_oracle_pins():
    """Multi-RHS PCG: one column is the single-RHS PCG bit for bit (same loop); with several columns,
    a column that needs the full iteration count also matches its single solve; every column solves."""
    p, i, v = st.laplace3d(9, 8, 7)
    n = len(p) - 1
    d, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, d)
    B = np.stack([st.uniform(90 + k, n) for k in range(3)])
    tol = 1e-10 * max(np.linalg.norm(b) for b in B)
    it1, X1, h1 = oracle.cg_batch(p, i, v, B[:1], tol=tol, prec=oracle.PREC_ILU, mdiag=d, mdata=f)
    it0, x0, h0 = oracle.cg(p, i, v, B[0], tol=tol, prec=oracle.PREC_ILU, mpntr=p, mindx=i, mdiag=d, mdata=f)
    assert it1 == it0 and np.array_equal(bits(X1[0]), bits(x0)) and np.array_equal(bits(h1), bits(h0))
    it, X, h = oracle.cg_batch(p, i, v, B, tol=tol, prec=oracle.PREC_ILU, mdiag=d, mdata=f)
    singles = [oracle.cg(p, i, v, B[k], tol=tol, prec=oracle.PREC_ILU, mpntr=p, mindx=i, mdiag=d, mdata=f)
               for k in range(3)]
    assert it == max(s[0] for s in singles)
    for k in range(3):
        if singles[k][0] == it:
            assert np.array_equal(bits(X[k]), bits(singles[k][1]))
    a = dense(p, i, v, n)
    for k in range(3):
        assert np.linalg.norm(np.linalg.solve(a, B[k]) - X[k]) < 1e-8 * np.linalg.norm(X[k])


def test_adjoint_gradient_finite_differences():
    m = 4
    d = st.diffusion3d(m, seed=5)
    dt = st.diffusion3d(m, seed=6)
    f = st.gaussian_source(m)
    p, i = d["pntr"], d["indx"]
    n = len(p) - 1
    cptr, cnnz, ccoef = d["cptr"], d["cnnz"], d["ccoef"]

    def vals_of(k):
        v = np.zeros(len(i))
        for c in range(len(k)):
            for e in range(cptr[c], cptr[c + 1]):
                v[cnnz[e]] += ccoef[e] * k[c]
        return v

    np.testing.assert_allclose(vals_of(d["kappa"]), d["vals"], rtol=1e-15)
    xt = np.linalg.solve(dense(p, i, vals_of(dt["kappa"]), n), f)

    def loss(k):
        x = np.linalg.solve(dense(p, i, vals_of(k), n), f)
        return 0.5 * np.sum((x - xt) ** 2)

    k0 = d["kappa"]
    v0 = vals_of(k0)
    x = np.linalg.solve(dense(p, i, v0, n), f)
    tp, ti, tv = oracle.transpose(n, n, p, i, v0)
    dg, _ = oracle.diag(p, i)
    fac = oracle.ilu0(p, i, v0, dg)
    _, lam, _ = oracle.cg(tp, ti, tv, x - xt, tol=1e-14 * np.linalg.norm(x - xt), prec=oracle.PREC_ILU_T, mpntr=p,
                          mindx=i, mdiag=dg, mdata=fac)
    g = oracle.grad_coef(cptr, cnnz, ccoef, oracle.grad_vals(p, i, lam, x))
    h = 1e-6
    for c in np.random.default_rng(0).choice(len(k0), 10, replace=False):
        kp, km = k0.copy(), k0.copy()
        kp[c] += h
        km[c] -= h
        fd = (loss(kp) - loss(km)) / (2 * h)
        assert abs(fd - g[c]) <= 1e-6 * np.abs(g).max()


def test_team_cg_matches_canonical_to_rounding():
    d = st.batch_laplace(3, seed=2, side_lo=14, side_hi=20)
    ro, p, i, v, dv = d["row_off"], d["pntr"], d["indx"], d["vals"], d["dvals"]
    b = st.uniform(9, len(dv))
    for s in range(3):
        r0, r1 = ro[s], ro[s + 1]
        z0, z1 = p[r0], p[r1]
        pp = p[r0:r1 + 1] - z0
        tol = 1e-10 * np.linalg.norm(b[r0:r1])
        it1, x1, _ = oracle.cg_team(pp, i[z0:z1], v[z0:z1], dv[r0:r1], b[r0:r1], tol, 256)
        it2, x2, _ = oracle.cg(pp, i[z0:z1], v[z0:z1], b[r0:r1], tol=tol, prec=oracle.PREC_JACOBI, mdata=dv[r0:r1])
        assert abs(it1 - it2) <= 1
        np.testing.assert_allclose(x1, x2, rtol=1e-8)


# ---------------------------------------------------------------- generators
def test_generator_sizes_and_structure():
    p, i, v = st.laplace3d(128, 128, 128)
    assert len(p) - 1 == 2097152 and len(i) == 14581760
    p, i, v = st.convdiff27(12, seed=3)
    assert len(i) == (3 * 12 - 2) ** 3
    A = dense(p, i, v, 12 ** 3)
    assert np.all(np.diag(A) >= np.abs(A).sum(1) - np.abs(np.diag(A)) - 1e-12)  # weakly diagonally dominant
    assert not np.allclose(A, A.T)  # convection makes it nonsymmetric
    d = st.diffusion3d(6, seed=5)
    A = dense(d["pntr"], d["indx"], d["vals"], 216)
    assert np.allclose(A, A.T)
    assert np.linalg.eigvalsh(A).min() > 0
    bl = st.batch_laplace(4096, seed=4)
    sides = np.diff(bl["row_off"])
    assert len(sides) == 4096 and sides.min() >= 196 and sides.max() <= 4900
    assert np.array_equal(st.uniform(3, 10), st.uniform(3, 10))


def test_threaded_oracle_is_bit_identical():
    # the threaded oracle splits only order-free work (rows, canonical blocks, rows of a level)
    p, i, v = st.laplace3d(24, 20, 18)
    n = len(p) - 1
    b = st.uniform(21, n)
    d, _ = oracle.diag(p, i)
    q, j, w = st.convdiff27(18, seed=5)  # n = 5832: above the threading threshold
    m = len(q) - 1
    c = st.uniform(22, m)
    dq, _ = oracle.diag(q, j)

    def run():
        f = oracle.ilu0(p, i, v, d)
        out = [f, oracle.spmv("N", n, n, 1.5, p, i, v, b, -0.5, b), np.array([oracle.cdot(b, b[::-1].copy())])]
        out += list(oracle.cg(p, i, v, b, tol=1e-10 * oracle.cnrm2(b), prec=oracle.PREC_ILU, mpntr=p, mindx=i,
                              mdiag=d, mdata=f)[1:])
        fq = oracle.ilu0(q, j, w, dq)
        out += list(oracle.gmres(q, j, w, c, tol=1e-11, restart=20, dg=dq, ilu=fq)[1:])
        for kw in ({"fp32": True}, {"fp32": True, "jacobi": 3}, {"fp32": False}, {"fp32": False, "jacobi": 3}):
            out += list(oracle.gmres_ir(q, j, w, dq, fq, c, **kw)[1:])
        return out

    t0 = oracle.get_threads()
    try:
        oracle.set_threads(1)
        one = run()
        oracle.set_threads(4)
        four = run()
    finally:
        oracle.set_threads(t0)
    assert len(one) == len(four)
    for a, b_ in zip(one, four):
        assert np.array_equal(bits(np.asarray(a)), bits(np.asarray(b_)))
