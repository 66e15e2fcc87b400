"""hala_cg_f64_noalloc: the allocation-free CG entry point, through the C ABI,
against hala_cg_f64 on the same inputs (bit for bit), two calls in a row with
the second smaller, for each preconditioner; sizes beyond the scratch are
refused; and the call takes no device memory."""
import ctypes as C

import numpy as np
import pytest

import hala_amd as H
import oracle
from hala_amd import _lib
from hala_amd import stencils as st
from tests.conftest import bits

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 21


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _solve(fn, A, gv, prec, ilu, gd, gb, gx, tol, gh):
    it = C.c_int()
    rc = fn(A.handle, _ptr(gv), prec, ilu.handle if ilu is not None else C.c_void_p(0), _ptr(gd), _ptr(gb), _ptr(gx),
            tol, 100000, C.byref(it), _ptr(gh), int(gh.numel()))
    assert rc == 0, _lib.lib().hala_last_error()
    return it.value


def _case(engine, side, prec, seed):
    import torch

    p, i, v = st.laplace3d(side, side - 1, side + 1)
    n = len(p) - 1
    b = st.uniform(seed, n)
    gp, gi, gv, gb = engine.load(p), engine.load(i), engine.load(v), engine.load(b)
    A = H.make_sparse_matrix(engine, n, gp, gi, gv)
    ilu = H.make_ilu(engine, gp, gi, gv) if prec == 1 else None
    dg = engine.load(v[oracle.diag(p, i)[0]]) if prec == 2 else None
    tol = 1e-10 * oracle.cnrm2(b)
    L = _lib.lib()
    out = []
    for fn in (L.hala_cg_f64, L.hala_cg_f64_noalloc):
        gx = engine.vector(n)
        gh = torch.zeros(2048, dtype=torch.float64, device=gx.device)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        it = _solve(fn, A, gv, prec, ilu, dg, gb, gx, tol, gh)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        out.append((it, gx.cpu().numpy(), gh.cpu().numpy()[: it - 1], free0 - free1))
    return n, out


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_noalloc_matches_cg_two_calls(prec):
    engine = H.gpu_engine(0)  # fresh context: its CG buffers start empty
    print(f"[noalloc] prec {prec}", flush=True)
    for side, seed in ((48, 61), (20, 62)):  # the second call is smaller
        n, ((it, x, h, _), (it2, x2, h2, used)) = _case(engine, side, prec, seed)
        assert it2 == it and it > 2
        assert np.array_equal(bits(x2), bits(x))
        assert np.array_equal(bits(h2), bits(h))
        assert used <= 0, f"the allocation-free call took {used} bytes of device memory"


def test_noalloc_refuses_oversize():
    engine = H.gpu_engine(0)
    n = MAX_ROWS + 1
    # a diagonal pattern just past the limit: the entry point refuses it before any launch
    p = np.arange(n + 1, dtype=np.int32)
    i = np.arange(n, dtype=np.int32)
    gp, gi = engine.load(p), engine.load(i)
    gv = engine.vector(n, 1.0)
    A = H.make_sparse_matrix(engine, n, gp, gi, gv)
    gb, gx = engine.vector(n, 1.0), engine.vector(n)
    it = C.c_int()
    L = _lib.lib()
    rc = L.hala_cg_f64_noalloc(A.handle, _ptr(gv), 0, C.c_void_p(0), C.c_void_p(0), _ptr(gb), _ptr(gx), 0.0, 10,
                               C.byref(it), C.c_void_p(0), 0)
    assert rc == -1  # HALA_ERR_ARG
    assert b"exceed the scratch" in L.hala_last_error()
