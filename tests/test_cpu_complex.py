"""CPU pins of the oracle's complex restatements: small-integer data (every
sum and product exact, so any order gives the same bits) against numpy, and
the defining identities of ILU(0) against numpy algebra."""
import numpy as np
import pytest

import oracle
from hala_amd import stencils as st


def cplx_matrix(p, i, seed, ctype=np.complex128, integer=True):
    rng = np.random.default_rng(seed)
    k = i.size
    if integer:
        v = rng.integers(-4, 5, k) + 1j * rng.integers(-4, 5, k)
    else:
        v = rng.standard_normal(k) + 1j * rng.standard_normal(k)
    n = p.size - 1
    # dominant diagonal so ILU(0) pivots stay away from zero
    for r in range(n):
        for j in range(p[r], p[r + 1]):
            if i[j] == r:
                v[j] = 40.0 + 3.0j if integer else 12.0 + 1.5j
    return v.astype(ctype)


def dense(p, i, v, n, m):
    A = np.zeros((n, m), v.dtype)
    for r in range(n):
        for j in range(p[r], p[r + 1]):
            A[r, i[j]] += v[j]
    return A


@pytest.mark.parametrize("ctype", [np.complex128, np.complex64])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_zspmv_exact_on_integers(trans, ctype):
    p, i, _ = st.laplace2d(9, 7)
    n = p.size - 1
    v = cplx_matrix(p, i, 3, ctype)
    rng = np.random.default_rng(5)
    x = (rng.integers(-3, 4, n) + 1j * rng.integers(-3, 4, n)).astype(ctype)
    y0 = (rng.integers(-3, 4, n) + 1j * rng.integers(-3, 4, n)).astype(ctype)
    alpha, beta = 2.0 - 1.0j, -1.0 + 3.0j
    A = dense(p, i, v, n, n)
    op = {"N": A, "T": A.T, "C": A.conj().T}[trans]
    want = alpha * (op @ x) + beta * y0
    y = y0.copy()
    oracle.zspmv(trans, n, n, alpha, p, i, v, x, beta, y)
    assert np.array_equal(y, want.astype(ctype))
    y = y0.copy()
    oracle.zspmv(trans, n, n, alpha, p, i, v, x, 0j, y)
    assert np.array_equal(y, (alpha * (op @ x)).astype(ctype))


@pytest.mark.parametrize("ctype", [np.complex128, np.complex64])
def test_zdot_znrm2_exact_on_integers(ctype):
    rng = np.random.default_rng(2)
    for n in (1, 255, 256, 257, 5000):
        x = (rng.integers(-50, 51, n) + 1j * rng.integers(-50, 51, n)).astype(ctype)
        y = (rng.integers(-50, 51, n) + 1j * rng.integers(-50, 51, n)).astype(ctype)
        assert oracle.zdot(x, y) == complex(np.vdot(x, y))
        assert oracle.zdot(x, y, conj=False) == complex(np.dot(x, y))
        ss = float(np.sum(x.real.astype(np.float64) ** 2 + x.imag.astype(np.float64) ** 2))
        rt = np.float64 if ctype == np.complex128 else np.float32
        assert oracle.znrm2(x) == rt(np.sqrt(rt(ss)))


@pytest.mark.parametrize("ctype, tol", [(np.complex128, 1e-13), (np.complex64, 2e-5)])
def test_zilu0_identities(ctype, tol):
    # tridiagonal: ILU(0) is the exact LU (tests/solvers_tests.hpp:26), so the apply solves A r = x
    n = 40
    p = np.zeros(n + 1, np.int32)
    rows = []
    for r in range(n):
        cols = [c for c in (r - 1, r, r + 1) if 0 <= c < n]
        rows.append(cols)
        p[r + 1] = p[r] + len(cols)
    i = np.array([c for cols in rows for c in cols], np.int32)
    v = cplx_matrix(p, i, 8, ctype, integer=False)
    dg, _ = oracle.diag(p, i)
    f = oracle.zilu0(p, i, v, dg)
    A = dense(p, i, v, n, n).astype(np.complex128)
    x = (np.random.default_rng(1).standard_normal(n) + 1j).astype(ctype)
    r = oracle.zilu_apply(p, i, dg, f, x)
    assert np.linalg.norm(A @ r.astype(np.complex128) - x) <= tol * 50 * np.linalg.norm(x)
    # general pattern (2D 5-point): (L U)_ij = A_ij on the pattern, L unit lower, U upper
    p, i, _ = st.laplace2d(8, 6)
    n = p.size - 1
    v = cplx_matrix(p, i, 9, ctype, integer=False)
    dg, _ = oracle.diag(p, i)
    f = oracle.zilu0(p, i, v, dg).astype(np.complex128)
    L = np.eye(n, dtype=np.complex128)
    U = np.zeros((n, n), np.complex128)
    for r in range(n):
        for j in range(p[r], p[r + 1]):
            (L if i[j] < r else U)[r, i[j]] = f[j]
    LU = L @ U
    A = dense(p, i, v, n, n).astype(np.complex128)
    mask = dense(p, i, np.ones(i.size), n, n) != 0
    assert np.max(np.abs((LU - A)[mask])) <= tol * 100
    # the apply is (LU)^-1 x
    x = (np.random.default_rng(4).standard_normal(n) * (1 - 2j)).astype(ctype)
    r = oracle.zilu_apply(p, i, dg, f.astype(ctype), x)
    assert np.linalg.norm(LU @ r.astype(np.complex128) - x) <= tol * 100 * np.linalg.norm(x)
