"""Sparse matrix times dense matrix (sparse_gemm) on the GPU for the four
scalar types: every column of C bit-identical to the oracle's product of
that column (the CPU engine computes gemm column by column,
sparse/hala_sparse_utils.hpp:127-167), for op(A) in N, T, C and op(B) in
N, T, C, on a rectangular matrix, with beta zero (C not read) and not."""
import numpy as np
import pytest
import torch

import hala_amd as H
import oracle
from tests.conftest import bits

pytestmark = pytest.mark.gpu

TT = {np.float64: torch.float64, np.float32: torch.float32, np.complex128: torch.complex128,
      np.complex64: torch.complex64}


def rect(nr, nc, ct, seed):
    rng = np.random.default_rng(seed)
    p, i, v = [0], [], []
    for r in range(nr):
        cols = np.sort(rng.choice(nc, size=min(int(rng.integers(0, 6)), nc), replace=False))
        i.extend(cols)
        w = rng.uniform(-1, 1, cols.size)
        if np.issubdtype(ct, np.complexfloating):
            w = w + 1j * rng.uniform(-1, 1, cols.size)
        v.extend(w)
        p.append(len(i))
    return np.array(p, np.int32), np.array(i, np.int32), np.array(v).astype(ct)


def dev(a, ct=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    return t if ct is None else t.to(TT[ct])


def raw(a):
    a = np.ascontiguousarray(a)
    return bits(a.view(np.float64 if a.dtype in (np.complex128, np.float64) else np.float32))


def ref_mv(trans, nr, nc, alpha, p, i, v, x, beta, y):
    y = y.copy()
    if v.dtype == np.float64:
        return oracle.spmv(trans, nr, nc, alpha, p, i, v, x, beta, y)
    if v.dtype == np.float32:
        if trans == "N":
            return oracle.spmv_f32(p, i, v, x, alpha=alpha, beta=beta, y=y)
        # the complex64 product with zero imaginary parts is the FP32 real one term by term
        yc = y.astype(np.complex64)
        oracle.zspmv(trans, nr, nc, complex(alpha), p, i, v.astype(np.complex64), x.astype(np.complex64),
                     complex(beta), yc)
        return yc.real.astype(np.float32)
    oracle.zspmv(trans, nr, nc, alpha, p, i, v, x, beta, y)
    return y


@pytest.mark.parametrize("ct, alpha, beta", [(np.float64, 1.5, -0.5), (np.float32, 0.75, 2.0),
                                             (np.complex128, 0.5 - 1j, 0.25 + 0.5j), (np.complex64, 1.0 + 1j, 0.0)],
                         ids=["f64", "f32", "z", "c"])
@pytest.mark.parametrize("transa", ["N", "T", "C"])
@pytest.mark.parametrize("transb", ["N", "T", "C"])
def test_spmm_columns_are_spmv(engine, ct, alpha, beta, transa, transb):
    real = not np.issubdtype(ct, np.complexfloating)
    if real and "C" in (transa, transb):
        pytest.skip("'C' equals 'T' for real values")
    nr, nc, ncol = 37, 29, 4
    p, i, v = rect(nr, nc, ct, 1)
    M, K = (nr, nc) if transa == "N" else (nc, nr)
    rng = np.random.default_rng(2)

    def rnd(n):
        a = rng.uniform(-1, 1, n)
        return (a + 1j * rng.uniform(-1, 1, n) if not real else a).astype(ct)

    if transb == "N":
        b_rows, b_cols, ldb = K, ncol, K + 3
    else:
        b_rows, b_cols, ldb = ncol, K, ncol + 2
    B = rnd(ldb * b_cols)
    ldc = M + 1
    C0 = rnd(ldc * ncol)
    A = H.make_sparse_matrix(engine, nc, dev(p), dev(i), dev(v, ct))
    gC = dev(C0, ct)
    A.gemm(transa, transb, b_rows, b_cols, alpha, dev(B, ct), ldb, beta, gC, ldc)
    out = gC.cpu().numpy()
    for k in range(ncol):
        x = B[k * ldb:k * ldb + K] if transb == "N" else B[k::ldb][:K]
        if transb == "C":
            x = np.conj(x)
        want = ref_mv(transa, nr, nc, alpha, p, i, v, x, beta, C0[k * ldc:k * ldc + M])
        assert np.array_equal(raw(out[k * ldc:k * ldc + M]), raw(want)), k
        assert np.array_equal(raw(out[k * ldc + M:(k + 1) * ldc]), raw(C0[k * ldc + M:(k + 1) * ldc]))


def test_spmm_shape_mismatch_is_an_error(engine):
    p, i, v = rect(5, 4, np.float64, 3)
    A = H.make_sparse_matrix(engine, 4, dev(p), dev(i), dev(v, np.float64))
    B = torch.zeros(20, dtype=torch.float64, device="cuda")
    C = torch.zeros(20, dtype=torch.float64, device="cuda")
    with pytest.raises(H.HalaError, match="column count"):
        A.gemm("N", "N", 5, 2, 1.0, B, 5, 0.0, C, 5)
