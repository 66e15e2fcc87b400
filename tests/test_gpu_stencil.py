"""The pencil schedule of the ILU(0) sweeps on stencil patterns: bit for bit the
oracle's apply (the reference's apply_ilu_array order), in FP64, FP32 and FP32
between FP64 vectors, on grids whose sides are and are not multiples of the
8 x 8 (64 x 1 in 2-D) tiles, repeated (the sentinels re-arm themselves) and in
place; patterns that are not stencils keep the general sweeps.  The schedule
is opt-in (option tri_stencil = 1)."""
import numpy as np
import pytest

import hala_amd as H
import oracle
from hala_amd import stencils as st
from tests.conftest import bits

pytestmark = pytest.mark.gpu

KIND = {"5pt": 1, "9pt": 2, "7pt": 3, "27pt": 4}

CASES = {
    "lap3d_13x11x9": (lambda: st.laplace3d(13, 11, 9), "7pt", (13, 11, 9)),
    "lap3d_24": (lambda: st.laplace3d(24, 24, 24), "7pt", (24, 24, 24)),
    "lap3d_40x9x17": (lambda: st.laplace3d(40, 9, 17), "7pt", (40, 9, 17)),
    "lap3d_5x70x3": (lambda: st.laplace3d(5, 70, 3), "7pt", (5, 70, 3)),
    "cd27_9": (lambda: st.convdiff27(9, seed=3), "27pt", (9, 9, 9)),
    "cd27_17": (lambda: st.convdiff27(17, seed=4), "27pt", (17, 17, 17)),
    "lap2d_16x12": (lambda: st.laplace2d(16, 12), "5pt", (16, 12, 1)),
    "lap2d_300x70": (lambda: st.laplace2d(300, 70), "5pt", (300, 70, 1)),
    "pde_33": (lambda: st.discretize_pde(33)[:3], "5pt", (33, 33, 1)),
    "diff3d_10": (lambda: (lambda d: (d["pntr"], d["indx"], d["vals"]))(st.diffusion3d(10, seed=5)), "7pt",
                  (10, 10, 10)),
}


def dev(engine, a):
    return engine.load(a)


@pytest.fixture(autouse=True)
def pencil_sweeps(engine):
    # the pencil schedule is opt-in (option tri_stencil)
    saved = engine.get_option("tri_stencil")
    engine.set_option("tri_stencil", 1)
    yield
    engine.set_option("tri_stencil", saved)


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", sorted(CASES))
def test_stencil_apply_bitexact(engine, name):
    import torch

    gen, kind, grid = CASES[name]
    p, i, v = gen()
    n = len(p) - 1
    print(f"[stencil] {name} n={n}", flush=True)
    dg, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, dg)
    gp, gi, gv = dev(engine, p), dev(engine, i), dev(engine, v)
    M = H.make_ilu(engine, gp, gi, gv, want_f32=True)
    info = M.stencil_info()
    assert info["kind"] == KIND[kind] and (info["nx"], info["ny"], info["nz"]) == grid and info["in_use"], info
    for rep, seed in enumerate((71, 72, 73)):  # repeated applies: the sentinels re-arm themselves
        x = st.uniform(seed, n)
        gr = engine.vector(n)
        M.apply(dev(engine, x), gr)
        assert np.array_equal(bits(host(gr)), bits(oracle.ilu_apply(p, i, dg, f, x))), rep
        # FP32 factor between FP64 vectors, and pure FP32
        ref32 = oracle.ilu_apply_f32(p, i, dg, f.astype(np.float32), x.astype(np.float32))
        M.apply_f32x64(dev(engine, x), gr)
        assert np.array_equal(bits(host(gr)), bits(ref32.astype(np.float64))), rep
        g32 = engine.vector(n, dtype=torch.float32)
        M.apply(engine.load(x.astype(np.float32)), g32)
        assert np.array_equal(bits(host(g32)), bits(ref32)), rep
    # in place
    x = st.uniform(74, n)
    gx = dev(engine, x)
    M.apply(gx, gx)
    assert np.array_equal(bits(host(gx)), bits(oracle.ilu_apply(p, i, dg, f, x)))
    # the general sweeps give the same bits
    saved = engine.get_option("tri_stencil")
    try:
        engine.set_option("tri_stencil", 0)
        assert not M.stencil_info()["in_use"]
        gr2 = engine.vector(n)
        M.apply(dev(engine, x), gr2)
        assert np.array_equal(bits(host(gr2)), bits(host(gx)))
    finally:
        engine.set_option("tri_stencil", saved)


def test_stencil_refactor_and_pcg(engine):
    # a refactorisation reloads the packed values; ILU-PCG through the pencil sweeps stays bit-exact
    import torch

    d = st.diffusion3d(16, seed=5)
    p, i = d["pntr"], d["indx"]
    n = len(p) - 1
    gp, gi = dev(engine, p), dev(engine, i)
    M = None
    for seed in (5, 6):
        v = st.diffusion3d(16, seed=seed)["vals"]
        gv = dev(engine, v)
        if M is None:
            M = H.make_ilu(engine, gp, gi, gv)
        else:
            M.factor(gv)
        assert M.stencil_info()["in_use"]
        b = st.uniform(seed + 80, n)
        tol = 1e-10 * oracle.cnrm2(b)
        dg, _ = oracle.diag(p, i)
        f = oracle.ilu0(p, i, v, dg)
        it, x, hist = oracle.cg(p, i, v, b, tol=tol, prec=oracle.PREC_ILU, mpntr=p, mindx=i, mdiag=dg, mdata=f)
        gx = engine.vector(n)
        gh = torch.zeros(512, dtype=torch.float64, device=gx.device)
        git = H.solve_cg_ilu(engine, tol, gp, gi, gv, M, dev(engine, b), gx, hist=gh)
        assert git == it
        assert np.array_equal(bits(host(gh)[: it - 1]), bits(hist))
        assert np.array_equal(bits(host(gx)), bits(x))


def test_non_stencil_patterns_keep_general_sweeps(engine):
    # a random pattern, and a periodic 2-D Laplacian (wrap-around couplings leave the box)
    rng = np.random.default_rng(3)
    n = 3000
    rows = [np.unique(np.concatenate([[r], rng.choice(n, 6, replace=False)])) for r in range(n)]
    p = np.zeros(n + 1, np.int32)
    p[1:] = np.cumsum([len(r) for r in rows])
    i = np.concatenate(rows).astype(np.int32)
    v = np.where(i == np.repeat(np.arange(n), np.diff(p)), 20.0, -1.0)
    M = H.make_ilu(engine, dev(engine, p), dev(engine, i), dev(engine, v))
    assert M.stencil_info()["kind"] == 0
    nx = ny = 40
    cols, ptr = [], [0]
    for r in range(nx * ny):
        a, b_ = r % nx, r // nx
        c = {r, (a + 1) % nx + nx * b_, (a - 1) % nx + nx * b_, a + nx * ((b_ + 1) % ny), a + nx * ((b_ - 1) % ny)}
        cols.extend(sorted(c))
        ptr.append(len(cols))
    p, i = np.array(ptr, np.int32), np.array(cols, np.int32)
    v = np.where(i == np.repeat(np.arange(nx * ny), np.diff(p)), 4.5, -1.0)
    M = H.make_ilu(engine, dev(engine, p), dev(engine, i), dev(engine, v))
    assert M.stencil_info()["kind"] == 0
    x = st.uniform(9, nx * ny)
    gr = engine.vector(nx * ny)
    M.apply(dev(engine, x), gr)
    dg, _ = oracle.diag(p, i)
    assert np.array_equal(bits(host(gr)), bits(oracle.ilu_apply(p, i, dg, oracle.ilu0(p, i, v, dg), x)))
