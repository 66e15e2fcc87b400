"""Standalone triangular solves (sparse_trsv / sparse_trsm) for FP32 and
complex values on the GPU, bit-exact against the oracle's restatement of the
reference's sparse_trsv_array on the same inputs."""
import numpy as np
import pytest
import torch

import hala_amd as H
import oracle
from hala_amd import stencils as st
from tests.conftest import bits

pytestmark = pytest.mark.gpu


def tri3d(uplo, m=12, ct=np.float64, seed=0):
    p, i, v = st.laplace3d(m, m, m)
    n = len(p) - 1
    rng = np.random.default_rng(seed)
    tp, ti, tv = [0], [], []
    for r in range(n):
        c = i[p[r]:p[r + 1]]
        keep = c[c <= r] if uplo == "L" else c[c >= r]
        w = rng.uniform(-1.0, 1.0, keep.size)
        if np.issubdtype(ct, np.complexfloating):
            w = w + 1j * rng.uniform(-1.0, 1.0, keep.size)
        w[keep == r] = 5.0 + (0.5j if np.issubdtype(ct, np.complexfloating) else 0.0)
        ti.extend(keep)
        tv.extend(w)
        tp.append(len(ti))
    return np.array(tp, np.int32), np.array(ti, np.int32), np.array(tv).astype(ct)


TT = {np.float32: torch.float32, np.complex128: torch.complex128, np.complex64: torch.complex64,
      np.float64: torch.float64}


def to_dev(a, ct):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda").to(TT[ct])


def raw(a):
    a = np.ascontiguousarray(a)
    return bits(a.view(np.float64 if a.dtype in (np.complex128, np.float64) else np.float32))


def ref_solve(uplo, trans, dg, alpha, p, i, v, b):
    if v.dtype == np.float32:
        return oracle.trsv_f32(uplo, "T" if trans == "C" else trans, dg, alpha, p, i, v, b)
    if v.dtype == np.float64:
        return oracle.trsv(uplo, "T" if trans == "C" else trans, dg, alpha, p, i, v, b)
    return oracle.ztrsv(uplo, trans, dg, alpha, p, i, v, b)


CASES = [(np.float32, 1.5), (np.complex128, 0.75 - 0.5j), (np.complex64, 0.75 - 0.5j)]


@pytest.mark.parametrize("ct, alpha", CASES, ids=["f32", "z", "c"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("dg", ["N", "U"])
def test_trsv_types_bitexact(engine, ct, alpha, uplo, trans, dg):
    p, i, v = tri3d(uplo, ct=ct, seed=1)
    n = len(p) - 1
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, n)
    if np.issubdtype(ct, np.complexfloating):
        b = b + 1j * rng.uniform(-1, 1, n)
    b = b.astype(ct)
    gp, gi = to_dev(p, np.float64).to(torch.int32), to_dev(i, np.float64).to(torch.int32)
    T = H.make_triangular_matrix(engine, uplo, dg, gp, gi, to_dev(v, ct))
    gx = torch.zeros(n, dtype=TT[ct], device="cuda")
    want = ref_solve(uplo, trans, dg, alpha, p, i, v, b)
    for _ in range(2):  # repeated solves re-arm the plan's polled outputs
        T.trsv(trans, alpha, to_dev(b, ct), gx)
        assert np.array_equal(raw(gx.cpu().numpy()), raw(want))


@pytest.mark.parametrize("ct, alpha", [(np.float64, -0.5), (np.float32, 1.5), (np.complex128, 0.75 - 0.5j),
                                       (np.complex64, 2.0 + 1.0j)], ids=["f64", "f32", "z", "c"])
@pytest.mark.parametrize("transb", ["N", "T"])
def test_trsm_columns_match_trsv(engine, ct, alpha, transb):
    uplo, transa, nrhs = "L", "T", 3
    p, i, v = tri3d(uplo, m=9, ct=ct, seed=3)
    n = len(p) - 1
    rng = np.random.default_rng(4)
    cols = [rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if np.issubdtype(ct, np.complexfloating) else 0)
            for _ in range(nrhs)]
    cols = [c.astype(ct) for c in cols]
    if transb == "N":
        ldb = n + 5
        Bh = np.zeros(ldb * nrhs, ct)
        for k in range(nrhs):
            Bh[k * ldb:k * ldb + n] = cols[k]
    else:
        ldb = nrhs + 2
        Bh = np.zeros(ldb * n, ct)
        for k in range(nrhs):
            Bh[k::ldb][:n] = cols[k]
    gp, gi = to_dev(p, np.float64).to(torch.int32), to_dev(i, np.float64).to(torch.int32)
    T = H.make_triangular_matrix(engine, uplo, "N", gp, gi, to_dev(v, ct))
    gB = to_dev(Bh, ct)
    T.trsm(transa, transb, nrhs, alpha, gB, ldb)
    out = gB.cpu().numpy()
    for k in range(nrhs):
        want = ref_solve(uplo, transa, "N", alpha, p, i, v, cols[k])
        got = out[k * ldb:k * ldb + n] if transb == "N" else out[k::ldb][:n]
        assert np.array_equal(raw(got), raw(want)), k
    # entries outside the columns are untouched
    mask = np.ones(Bh.size, bool)
    for k in range(nrhs):
        if transb == "N":
            mask[k * ldb:k * ldb + n] = False
        else:
            mask[np.arange(n) * ldb + k] = False
    assert np.array_equal(raw(out[mask]), raw(Bh[mask]))


def test_trsm_rejects_short_ldb(engine):
    p, i, v = tri3d("L", m=4, ct=np.float64)
    n = len(p) - 1
    gp, gi = to_dev(p, np.float64).to(torch.int32), to_dev(i, np.float64).to(torch.int32)
    T = H.make_triangular_matrix(engine, "L", "N", gp, gi, to_dev(v, np.float64))
    gB = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    with pytest.raises(H.HalaError, match="ldb"):
        T.trsm("N", "N", 2, 1.0, gB, n - 1)
