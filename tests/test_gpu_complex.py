"""GPU tests of the complex scalar path (std::complex<double> "z" and <float>
"c"): BLAS-1, the CSR product and ILU(0) factor + apply, each bit-exact
against the oracle's restatement on the same inputs."""
import ctypes as C

import numpy as np
import pytest
import torch

import hala_amd as H
import oracle
from hala_amd import stencils as st
from hala_amd._lib import call
from tests.conftest import bits

pytestmark = pytest.mark.gpu

TYPES = [("z", np.complex128, torch.complex128, C.c_double, np.float64),
         ("c", np.complex64, torch.complex64, C.c_float, np.float32)]


def ptr(t):
    return C.c_void_p(t.data_ptr())


def dev(a, tt):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda").to(tt)


def host(t):
    return t.cpu().numpy()


def cbits(a):
    return bits(np.ascontiguousarray(a).view(np.float64 if a.dtype == np.complex128 else np.float32))


def cmatrix(p, i, seed, ctype):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(i.size) + 1j * rng.standard_normal(i.size)
    for r in range(p.size - 1):
        for j in range(p[r], p[r + 1]):
            if i[j] == r:
                v[j] = 9.0 + 2.0j
    return v.astype(ctype)


@pytest.mark.parametrize("sf, nt, tt, cr, rt", TYPES, ids=["z", "c"])
def test_complex_blas1_bitexact(engine, sf, nt, tt, cr, rt):
    rng = np.random.default_rng(11)
    for n in (1, 300, 70001):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(nt)
        y = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(nt)
        gx, gy = dev(x, tt), dev(y, tt)
        res = (cr * 2)()
        call("hala_dotc_" + sf, engine.handle, n, ptr(gx), ptr(gy), res)
        want = oracle.zdot(x, y)
        assert (rt(res[0]), rt(res[1])) == (rt(want.real), rt(want.imag))
        call("hala_dotu_" + sf, engine.handle, n, ptr(gx), ptr(gy), res)
        want = oracle.zdot(x, y, conj=False)
        assert (rt(res[0]), rt(res[1])) == (rt(want.real), rt(want.imag))
        nr = cr()
        call("hala_nrm2_" + sf, engine.handle, n, ptr(gx), C.byref(nr))
        assert rt(nr.value) == rt(oracle.znrm2(x))
        # y + a x and a x, elementwise with the documented product
        a = 0.75 - 1.25j
        call("hala_axpy_" + sf, engine.handle, n, a.real, a.imag, ptr(gx), ptr(gy))
        ar, ai = rt(a.real), rt(a.imag)
        xr, xi, yr, yi = x.real, x.imag, y.real, y.imag
        want = (yr + (ar * xr - ai * xi)) + 1j * (yi + (ar * xi + ai * xr))
        assert np.array_equal(cbits(host(gy)), cbits(want.astype(nt)))
        call("hala_scal_" + sf, engine.handle, n, a.real, a.imag, ptr(gx))
        want = (ar * xr - ai * xi) + 1j * (ar * xi + ai * xr)
        assert np.array_equal(cbits(host(gx)), cbits(want.astype(nt)))


@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("sf, nt, tt, cr, rt", TYPES, ids=["z", "c"])
def test_complex_spmv_bitexact(engine, sf, nt, tt, cr, rt, trans):
    p, i, _ = st.laplace3d(12, 10, 9)
    n = p.size - 1
    v = cmatrix(p, i, 2, nt)
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(nt)
    y0 = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(nt)
    A = H.make_sparse_matrix(engine, n, dev(p, torch.int32), dev(i, torch.int32), dev(v, tt))
    for alpha, beta in ((1.0 + 0j, 0j), (0.5 - 2.0j, -1.0 + 0.25j)):
        gy = dev(y0, tt)
        call("hala_spmv_" + sf, A.handle, trans.encode(), alpha.real, alpha.imag, ptr(A.vals), ptr(dev(x, tt)),
             beta.real, beta.imag, ptr(gy))
        want = y0.copy()
        oracle.zspmv(trans, n, n, alpha, p, i, v, x, beta, want)
        assert np.array_equal(cbits(host(gy)), cbits(want))


@pytest.mark.parametrize("sf, nt, tt, cr, rt", TYPES, ids=["z", "c"])
def test_complex_ilu_bitexact(engine, sf, nt, tt, cr, rt):
    for gen in (lambda: st.laplace3d(14, 12, 10), lambda: st.convdiff27(10, seed=2)):
        p, i, _ = gen()
        n = p.size - 1
        v = cmatrix(p, i, 7, nt)
        gp, gi = dev(p, torch.int32), dev(i, torch.int32)
        M = H.gpu_ilu(engine, gp, gi, None)
        call("hala_ilu_factor_" + sf, M.handle, ptr(dev(v, tt)))
        x = (np.random.default_rng(5).standard_normal(n) * (1 + 0.5j)).astype(nt)
        gr = torch.zeros(n, dtype=tt, device="cuda")
        call("hala_ilu_apply_" + sf, M.handle, ptr(dev(x, tt)), ptr(gr))
        dg, _ = oracle.diag(p, i)
        f = oracle.zilu0(p, i, v, dg)
        assert np.array_equal(cbits(host(gr)), cbits(oracle.zilu_apply(p, i, dg, f, x)))


def test_complex_ilu_apply_before_factor_is_an_error(engine):
    p, i, _ = st.laplace2d(6, 6)
    M = H.gpu_ilu(engine, dev(p, torch.int32), dev(i, torch.int32), None)
    g = torch.zeros(p.size - 1, dtype=torch.complex128, device="cuda")
    with pytest.raises(H.HalaError, match="before factorisation"):
        call("hala_ilu_apply_z", M.handle, ptr(g), ptr(g))
