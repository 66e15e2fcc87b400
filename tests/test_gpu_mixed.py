"""GPU tests of the mixed engine (host numpy arrays in and out, the work on
the device): every operation equals the gpu_engine path on the same data bit
for bit, and outputs land in the caller's arrays."""
import numpy as np
import pytest
import torch

import hala_amd as H
from hala_amd import stencils as st
from tests.conftest import bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    p, i, v = st.laplace3d(24, 20, 16)
    n = p.size - 1
    return p, i, v, st.uniform(5, n), st.uniform(6, n)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_mixed_blas_and_product(engine, problem):
    p, i, v, x, y = problem
    me = H.mixed_engine(engine)
    assert H.dot(me, x, y) == H.dot(engine, dev(x), dev(y))
    assert H.norm2(me, x) == H.norm2(engine, dev(x))
    ym, yg = y.copy(), dev(y)
    H.axpy(me, -0.75, x, ym)
    H.axpy(engine, -0.75, dev(x), yg)
    assert np.array_equal(bits(ym), bits(yg.cpu().numpy()))
    H.scal(me, 3.0, ym)
    H.scal(engine, 3.0, yg)
    assert np.array_equal(bits(ym), bits(yg.cpu().numpy()))
    z = np.zeros_like(x)
    H.vcopy(me, x, z)
    assert np.array_equal(bits(z), bits(x))
    n = p.size - 1
    om, og = np.zeros(n), torch.zeros(n, dtype=torch.float64, device="cuda")
    H.sparse_gemv(me, "N", n, n, 2.0, p, i, v, x, 0.0, om)
    H.sparse_gemv(engine, "N", n, n, 2.0, dev(p), dev(i), dev(v), dev(x), 0.0, og)
    assert np.array_equal(bits(om), bits(og.cpu().numpy()))


def test_mixed_ilu_and_solvers(engine, problem):
    p, i, v, b, x0 = problem
    n = p.size - 1
    me = H.mixed_engine(engine)
    milu = H.make_ilu(me, p, i, v)
    gp, gi, gv = dev(p), dev(i), dev(v)
    gilu = H.make_ilu(engine, gp, gi, gv)
    r = np.zeros(n)
    milu.apply(b, r)
    gr = engine.vector(n)
    gilu.apply(dev(b), gr)
    assert np.array_equal(bits(r), bits(gr.cpu().numpy()))
    # PCG: the host x receives the device result; identical to the gpu_engine solve
    x = np.zeros(n)
    it = H.solve_cg_ilu(me, 1e-9, p, i, v, milu, b, x)
    gx = engine.vector(n)
    itg = H.solve_cg_ilu(engine, 1e-9, gp, gi, gv, gilu, dev(b), gx)
    assert it == itg and np.array_equal(bits(x), bits(gx.cpu().numpy()))
    # factorisation behind the scene
    x2 = np.zeros(n)
    assert H.solve_cg_ilu(me, 1e-9, p, i, v, None, b, x2) == it and np.array_equal(bits(x2), bits(x))
    # GMRES(10) with ILU
    x3 = np.zeros(n)
    tot, _ = H.solve_gmres_ilu(me, 1e-9, 10, p, i, v, milu, b, x3)
    gx3 = engine.vector(n)
    totg, _ = H.solve_gmres_ilu(engine, 1e-9, 10, gp, gi, gv, gilu, dev(b), gx3)
    assert tot == totg and np.array_equal(bits(x3), bits(gx3.cpu().numpy()))
    # three right-hand sides with one shared matrix
    B = np.concatenate([b, x0, b - x0])
    X = np.zeros(3 * n)
    itb = H.solve_batch_cg_ilu(me, 1e-9, p, i, v, milu, B, X)
    GX = engine.vector(3 * n)
    itbg = H.solve_batch_cg_ilu(engine, 1e-9, gp, gi, gv, gilu, dev(B), GX)
    assert itb == itbg and np.array_equal(bits(X), bits(GX.cpu().numpy()))
