"""The oracle's float and complex triangular solves, pinned independently of
their own formulas: a unit-diagonal complex solve with zero imaginary parts is
bit-identical to the real FP64 solve (complex products then reduce to real
ones exactly); every solve satisfies op(T) x = alpha b against a dense
product (catches a transposed operand, a missing conjugate or a wrong
diagonal); 'C' on T equals 'T' on conj(T) bit for bit; the FP32 solve is
bit-identical to the FP64 one on small-integer data where both are exact."""
import numpy as np
import pytest

import oracle
from hala_amd import stencils as st


def triangle(uplo, m=7, seed=0, cplx=None):
    """One triangle of a 2-D 5-point pattern, the diagonal last (L) or first (U)."""
    p, i, _ = st.laplace2d(m, m)
    n = len(p) - 1
    rng = np.random.default_rng(seed)
    tp, ti, tv = [0], [], []
    for r in range(n):
        c = i[p[r]:p[r + 1]]
        keep = c[c <= r] if uplo == "L" else c[c >= r]
        vals = rng.uniform(-1.0, 1.0, keep.size)
        if cplx is not None:
            vals = vals + 1j * rng.uniform(-1.0, 1.0, keep.size)
        vals[keep == r] = 4.0 + (1.5j if cplx is not None else 0.0)
        ti.extend(keep)
        tv.extend(vals)
        tp.append(len(ti))
    dt = cplx if cplx is not None else np.float64
    return np.array(tp, np.int32), np.array(ti, np.int32), np.array(tv, dt)


def dense(p, i, v):
    n = len(p) - 1
    A = np.zeros((n, n), v.dtype)
    for r in range(n):
        A[r, i[p[r]:p[r + 1]]] = v[p[r]:p[r + 1]]
    return A


def op(A, trans):
    return A if trans == "N" else (A.T if trans == "T" else A.conj().T)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_unit_complex_with_real_data_equals_real(uplo, trans):
    p, i, v = triangle(uplo, seed=1)
    n = len(p) - 1
    b = st.uniform(2, n)
    want = oracle.trsv(uplo, trans, "U", 0.75, p, i, v, b)
    got = oracle.ztrsv(uplo, trans, "U", 0.75, p, i, v.astype(np.complex128), b.astype(np.complex128))
    assert np.array_equal(got.real.view(np.int64), want.view(np.int64))
    assert np.all(got.imag == 0)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("dg", ["N", "U"])
@pytest.mark.parametrize("ct, tol", [(np.complex128, 1e-12), (np.complex64, 2e-5)], ids=["z", "c"])
def test_complex_solves_dense_system(uplo, trans, dg, ct, tol):
    p, i, v = triangle(uplo, seed=3, cplx=ct)
    n = len(p) - 1
    rng = np.random.default_rng(4)
    b = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(ct)
    alpha = 0.5 - 0.25j
    x = oracle.ztrsv(uplo, trans, dg, alpha, p, i, v, b)
    A = dense(p, i, v.astype(np.complex128))
    if dg == "U":
        np.fill_diagonal(A, 1.0)
    r = op(A, trans) @ x.astype(np.complex128) - alpha * b.astype(np.complex128)
    assert np.linalg.norm(r) <= tol * np.linalg.norm(alpha * b)


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_conjugate_transpose_is_transpose_of_conjugate(uplo):
    p, i, v = triangle(uplo, seed=5, cplx=np.complex128)
    n = len(p) - 1
    b = (st.uniform(6, n) + 1j * st.uniform(7, n)).astype(np.complex128)
    xc = oracle.ztrsv(uplo, "C", "N", 1.25 + 0.5j, p, i, v, b)
    xt = oracle.ztrsv(uplo, "T", "N", 1.25 + 0.5j, p, i, np.conj(v), b)
    assert np.array_equal(xc.view(np.int64), xt.view(np.int64))


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_f32_exact_small_integers_equals_f64(uplo, trans):
    # unit diagonal, entries in {-1, 0, 1}, b small integers: every partial sum is
    # a small integer, exact in both precisions
    p, i, v = triangle(uplo, m=5, seed=8)
    v = np.sign(np.round(v))
    n = len(p) - 1
    b = np.arange(n, dtype=np.float64) % 5 - 2
    x64 = oracle.trsv(uplo, trans, "U", 1.0, p, i, v, b)
    assert np.all(np.abs(x64) < 2**20)
    x32 = oracle.trsv_f32(uplo, trans, "U", 1.0, p, i, v.astype(np.float32), b.astype(np.float32))
    assert np.array_equal(x32.astype(np.float64), x64)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_f32_solves_dense_system(uplo, trans):
    p, i, v = triangle(uplo, seed=9)
    n = len(p) - 1
    b = st.uniform(10, n).astype(np.float32)
    x = oracle.trsv_f32(uplo, trans, "N", 1.5, p, i, v.astype(np.float32), b)
    A = dense(p, i, v.astype(np.float32).astype(np.float64))
    r = op(A, trans) @ x.astype(np.float64) - 1.5 * b.astype(np.float64)
    assert np.linalg.norm(r) <= 1e-5 * np.linalg.norm(1.5 * b)
