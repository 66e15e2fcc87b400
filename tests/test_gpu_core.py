"""GPU parity of the core kernels against the CPU oracle (bit for bit).

Every FP64 kernel here follows the oracle's operation order exactly (no fused
multiply-add, sequential row sums, canonical reductions), so the comparisons
are on bit patterns unless a test says otherwise.
"""
import numpy as np
import pytest

import hala_amd as H
import oracle
from hala_amd import stencils as st
from tests.conftest import bits

pytestmark = pytest.mark.gpu


def dev(engine, a, dtype=None):
    return engine.load(a, dtype)


def host(t):
    return t.detach().cpu().numpy()


def upload(engine, p, i, v):
    return dev(engine, p), dev(engine, i), dev(engine, v)


# ---------------------------------------------------------------- BLAS-1
@pytest.mark.parametrize("n", [1, 7, 255, 256, 257, 1000, 65536 + 7, 300001])
def test_dot_nrm2_canonical(engine, n):
    x = st.uniform(11, n)
    y = st.uniform(12, n)
    gx, gy = dev(engine, x), dev(engine, y)
    assert bits(np.array([H.dot(engine, gx, gy)])) == bits(np.array([oracle.cdot(x, y)]))
    assert bits(np.array([H.norm2(engine, gx)])) == bits(np.array([oracle.cnrm2(x)]))


def test_sqrt_correctly_rounded(engine):
    # nrm2 = sqrt(cdot): single-element vectors make cdot exact (x*x), isolating the device sqrt
    rng = np.random.default_rng(0)
    for _ in range(200):
        v = np.array([rng.uniform(1e-3, 1e3) * 10.0 ** rng.integers(-30, 30)])
        g = dev(engine, v)
        assert H.norm2(engine, g) == np.sqrt(v[0] * v[0])


def test_axpy_scal_copy(engine):
    n = 100003
    x, y = st.uniform(1, n), st.uniform(2, n)
    gx, gy = dev(engine, x), dev(engine, y)
    H.axpy(engine, -0.37, gx, gy)
    assert np.array_equal(bits(host(gy)), bits(y + (-0.37) * x))
    H.scal(engine, 1.9, gy)
    assert np.array_equal(bits(host(gy)), bits(1.9 * (y + (-0.37) * x)))
    gz = engine.vector(n)
    H.vcopy(engine, gx, gz)
    assert np.array_equal(host(gz), x)


# ---------------------------------------------------------------- SpMV
MATS = {
    "lap2d_256": lambda: st.laplace2d(256, 256),
    "lap3d_40": lambda: st.laplace3d(40, 40, 40),
    "cd27_24": lambda: st.convdiff27(24, seed=3),
    "pde_33": lambda: st.discretize_pde(33)[:3],
}


@pytest.mark.parametrize("name", sorted(MATS))
def test_spmv_f64_bitexact(engine, name):
    p, i, v = MATS[name]()
    n = len(p) - 1
    gp, gi, gv = upload(engine, p, i, v)
    A = H.make_sparse_matrix(engine, n, gp, gi, gv)
    assert A.info()["kernel"] == 0
    x, y0 = st.uniform(3, n), st.uniform(4, n)
    for alpha, beta in [(1.0, 0.0), (2.5, -0.5), (-1.0, 1.0)]:
        gx, gy = dev(engine, x), dev(engine, y0)
        A.gemv("N", alpha, gx, beta, gy)
        ref = oracle.spmv("N", n, n, alpha, p, i, v, x, beta, y0)
        assert np.array_equal(bits(host(gy)), bits(ref)), (alpha, beta)
    # transposed product: source-row order of the reference scatter
    gx, gy = dev(engine, x), dev(engine, y0)
    A.gemv("T", 1.5, gx, 0.25, gy)
    ref = oracle.spmv("T", n, n, 1.5, p, i, v, x, 0.25, y0)
    assert np.array_equal(bits(host(gy)), bits(ref))


@pytest.mark.parametrize("hi,lo,short", [(32767, 32768, True), (32768, 32768, False), (32767, 32769, False)],
                         ids=["int16-limits", "past-high", "past-low"])
def test_spmv_block_offsets(engine, hi, lo, short):
    # columns streamed as int16 offsets from the 256-row block's first row when
    # every offset fits (-32768 .. 32767), the caller's int32 columns otherwise;
    # both match the oracle bit for bit, also inside CG
    n = 70000
    rows = []
    for r in range(n):
        r0 = r // 256 * 256
        c = {r}
        if r0 + hi < n and r % 7 == 0:
            c.add(r0 + hi)  # the largest offset above the block start
        if r0 - lo >= 0 and r % 5 == 0:
            c.add(r0 - lo)  # the largest below
        if r > 0:
            c.add(r - 1)
        rows.append(sorted(c))
    p = np.zeros(n + 1, np.int32)
    p[1:] = np.cumsum([len(c) for c in rows])
    i = np.concatenate([np.array(c, np.int32) for c in rows])
    v = np.random.default_rng(12).uniform(-1, 1, len(i))
    gp, gi, gv = upload(engine, p, i, v)
    A = H.make_sparse_matrix(engine, n, gp, gi, gv)
    assert A.info()["kernel"] == 0
    assert A.info()["index_bytes"] == (2 if short else 4)
    x, y0 = st.uniform(31, n), st.uniform(32, n)
    for alpha, beta in [(1.0, 0.0), (2.5, -0.5)]:
        gx, gy = dev(engine, x), dev(engine, y0)
        A.gemv("N", alpha, gx, beta, gy)
        ref = oracle.spmv("N", n, n, alpha, p, i, v, x, beta, y0)
        assert np.array_equal(bits(host(gy)), bits(ref)), (alpha, beta)


def test_spmv_block_offsets_stencils(engine):
    # the stencil operators (C2's 7-point 128^3 included) take the int16 offsets
    for p, i, v in (st.laplace3d(128, 128, 128), st.convdiff27(24, seed=2), st.laplace2d(300, 200)):
        n = len(p) - 1
        gp, gi, gv = upload(engine, p, i, v)
        A = H.make_sparse_matrix(engine, n, gp, gi, gv)
        assert A.info()["index_bytes"] == 2
        x = st.uniform(41, n)
        gx, gy = dev(engine, x), engine.vector(n)
        A.gemv("N", 1.0, gx, 0.0, gy)
        assert np.array_equal(bits(host(gy)), bits(oracle.spmv("N", n, n, 1.0, p, i, v, x, 0.0, np.zeros(n))))


def _long_row_pattern(n, per_row, seed, one_long=0):
    rng = np.random.default_rng(seed)
    rows = [np.unique(np.concatenate([[r], rng.choice(n, per_row, replace=False)])) for r in range(n)]
    if one_long:  # one row longer than 8192 entries (a ragged last chunk of the 64-wide chain)
        rows[n // 2] = np.arange(one_long, dtype=np.int64)
    p = np.zeros(n + 1, np.int32)
    p[1:] = np.cumsum([len(r) for r in rows])
    i = np.concatenate(rows).astype(np.int32)
    return p, i, rng.uniform(-1, 1, len(i))


@pytest.mark.parametrize("case", ["mean150", "row9001"])
def test_spmv_wave_kernel(engine, case):
    # long rows take the wavefront-per-row kernel; it adds each row's products in CSR
    # order (sparse/hala_sparse_utils.hpp:106-108), so it matches the oracle bit for bit
    import torch

    if case == "mean150":
        n = 2000
        p, i, v = _long_row_pattern(n, 150, 5)
    else:
        n = 9001
        p, i, v = _long_row_pattern(n, 3, 6, one_long=9001)
    gp, gi, gv = upload(engine, p, i, v)
    A = H.make_sparse_matrix(engine, n, gp, gi, gv)
    assert A.info()["kernel"] == 1
    x = st.uniform(8, n)
    y0 = st.uniform(9, n)
    gx, gy = dev(engine, x), dev(engine, y0)
    A.gemv("N", 1.5, gx, 0.25, gy)
    ref = oracle.spmv("N", n, n, 1.5, p, i, v, x, 0.25, y0)
    assert np.array_equal(bits(host(gy)), bits(ref))
    # FP32 values and vectors: the same order in FP32
    v32 = v.astype(np.float32)
    A32 = H.make_sparse_matrix(engine, n, gp, gi, dev(engine, v32))
    gx32 = dev(engine, x.astype(np.float32))
    gy32 = engine.vector(n, dtype=torch.float32)
    A32.gemv("N", 1.0, gx32, 0.0, gy32)
    assert np.array_equal(bits(host(gy32)), bits(oracle.spmv_f32(p, i, v32, x.astype(np.float32))))


def test_spmv_f32_variants(engine):
    import torch

    p, i, v = st.convdiff27(16, seed=3)
    n = len(p) - 1
    v32 = v.astype(np.float32)
    x = st.uniform(9, n)
    gp, gi = dev(engine, p), dev(engine, i)
    gv32 = dev(engine, v32)
    A = H.make_sparse_matrix(engine, n, gp, gi, gv32)
    # pure FP32
    gx32 = dev(engine, x.astype(np.float32))
    gy32 = engine.vector(n, dtype=torch.float32)
    A.gemv("N", 1.0, gx32, 0.0, gy32)
    ref = oracle.spmv_f32(p, i, v32, x.astype(np.float32))
    assert np.array_equal(bits(host(gy32)), bits(ref))
    # FP32 values and accumulation, FP64 vectors
    gx, gy = dev(engine, x), engine.vector(n)
    A.gemv("N", 1.0, gx, 0.0, gy)
    ref64 = 1.0 * oracle.spmv_f32(p, i, v32, x.astype(np.float32)).astype(np.float64) + 0.0
    assert np.array_equal(bits(host(gy)), bits(ref64))


def test_reference_known_answers(engine):
    # tests/sparse_tests.hpp:165-191 / tests/cuda_sparse_tests.hpp:10-30: 5x5 tridiagonal, alpha = 2
    p = np.array([0, 2, 5, 8, 11, 13], np.int32)
    i = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4], np.int32)
    v = np.array([2, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 2], np.float64)
    x = np.array([1, 2, 3, 4, 5], np.float64)
    A = np.zeros((5, 5))
    for r in range(5):
        A[r, i[p[r]:p[r + 1]]] = v[p[r]:p[r + 1]]
    gp, gi, gv = upload(engine, p, i, v)
    gy = engine.vector(5)
    H.sparse_gemv(engine, "N", 5, 5, 2.0, gp, gi, gv, dev(engine, x), 0.0, gy)
    assert np.array_equal(host(gy), 2.0 * A @ x)
    H.sparse_gemv(engine, "T", 5, 5, 2.0, gp, gi, gv, dev(engine, x), 0.0, gy)
    assert np.array_equal(host(gy), 2.0 * A.T @ x)
    # cmake/post_install_test.sh: diagonal 3x3 -> {4, 10, 18}
    gp = dev(engine, np.array([0, 1, 2, 3], np.int32))
    gi = dev(engine, np.array([0, 1, 2], np.int32))
    gv = dev(engine, np.array([4.0, 5.0, 6.0]))
    gy = engine.vector(3)
    H.sparse_gemv(engine, "N", 3, 3, 1.0, gp, gi, gv, dev(engine, np.array([1.0, 2.0, 3.0])), 0.0, gy)
    assert list(host(gy)) == [4.0, 10.0, 18.0]


def test_pattern_validation(engine):
    p = np.array([0, 2, 3], np.int32)
    i = np.array([1, 0, 5], np.int32)  # unsorted row 0, out-of-range column in row 1
    with pytest.raises(H.HalaError):
        H.make_sparse_matrix(engine, 2, dev(engine, p), dev(engine, i), dev(engine, np.ones(3)))


# ---------------------------------------------------------------- ILU(0)
def arrow2d(m=16, seed=9):
    """m x m 5-point Laplacian plus a dense last row/column and a row/column of
    ~100 entries: rows longer than the numeric kernel stages in LDS (and with
    more elimination pairs than it holds) next to rows that fit."""
    import scipy.sparse as sp

    p, i, v = st.laplace2d(m, m)
    n = len(p) - 1
    A = sp.csr_matrix((v, i, p), shape=(n, n)).tolil()
    rng = np.random.default_rng(seed)
    mid = n // 2
    for c in range(n - 1):
        A[n - 1, c] = A[c, n - 1] = rng.uniform(-0.01, 0.01)
    for c in range(mid - 60, n - 1, 2):
        if c != mid:
            A[mid, c] = A[c, mid] = rng.uniform(-0.01, 0.01)
    A = A.tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


ILU_MATS = {
    "lap2d_64": lambda: st.laplace2d(64, 64),
    "lap3d_24": lambda: st.laplace3d(24, 24, 24),
    "cd27_16": lambda: st.convdiff27(16, seed=3),
    "pde_20": lambda: st.discretize_pde(20)[:3],
    "arrow_16": arrow2d,
}


@pytest.mark.parametrize("name", sorted(ILU_MATS))
def test_ilu_symbolic_numeric_apply(engine, name):
    import torch

    p, i, v = ILU_MATS[name]()
    n = len(p) - 1
    gp, gi, gv = upload(engine, p, i, v)
    M = H.make_ilu(engine, gp, gi, gv, want_f32=True)
    dg, bad = oracle.diag(p, i)
    assert bad == -1
    for upper in (False, True):
        d_, lev, lptr, perm = M.symbolic(upper)
        olev, olptr, operm = oracle.levels(p, i, dg, upper)
        assert np.array_equal(host(d_), dg)
        assert np.array_equal(host(lev), olev)
        assert np.array_equal(host(lptr), olptr)
        assert np.array_equal(host(perm), operm)
    f = oracle.ilu0(p, i, v, dg)
    assert np.array_equal(bits(host(M.get_factor())), bits(f))
    x = st.uniform(21, n)
    gr = engine.vector(n)
    for _ in range(3):  # repeated applies re-arm the ready flags
        M.apply(dev(engine, x), gr)
        assert np.array_equal(bits(host(gr)), bits(oracle.ilu_apply(p, i, dg, f, x)))
    # FP32 factor, FP32 sweep
    x32 = x.astype(np.float32)
    g32 = engine.vector(n, dtype=torch.float32)
    M.apply(dev(engine, x32), g32)
    assert np.array_equal(bits(host(g32)), bits(oracle.ilu_apply_f32(p, i, dg, f.astype(np.float32), x32)))
    # FP32 sweep between FP64 vectors
    M.apply_f32x64(dev(engine, x), gr)
    ref = oracle.ilu_apply_f32(p, i, dg, f.astype(np.float32), x32).astype(np.float64)
    assert np.array_equal(bits(host(gr)), bits(ref))
    # adjoint apply (LU)^-T
    for _ in range(2):
        M.apply_t(dev(engine, x), gr)
        assert np.array_equal(bits(host(gr)), bits(oracle.ilu_apply_t(p, i, dg, f, x)))


TRI_MODES = [  # (tri_col, tri_col_cols, tri_streams, tri_pool_waves)
    (0, 0, 1, 0), (0, 0, 2, 0), (0, 0, 4, 0), (0, 0, 4, 8), (0, 0, 3, 12), (0, 0, 1, 100000),
    (1, 0, 1, 0), (1, 4, 1, 0), (1, 16, 1, 0), (1, 64, 1, 0)]


@pytest.mark.parametrize("mode", TRI_MODES, ids=lambda m: f"col{m[0]}x{m[1]}_streams{m[2]}_pool{m[3]}")
def test_tri_launch_modes_bitexact(engine, mode):
    """Every way of running the triangular sweeps (level pool, column-blocked) gives the same bits."""
    keys = ("tri_col", "tri_col_cols", "tri_streams", "tri_pool_waves")
    saved = [engine.get_option(k) for k in keys]
    try:
        for k, val in zip(keys, mode):
            engine.set_option(k, val)
        for name in ("lap3d_24", "cd27_16"):
            p, i, v = ILU_MATS[name]()
            n = len(p) - 1
            M = H.make_ilu(engine, *upload(engine, p, i, v), want_f32=True)
            dg, _ = oracle.diag(p, i)
            f = oracle.ilu0(p, i, v, dg)
            x = st.uniform(22, n)
            gr = engine.vector(n)
            for _ in range(3):
                M.apply(dev(engine, x), gr)
                assert np.array_equal(bits(host(gr)), bits(oracle.ilu_apply(p, i, dg, f, x))), name
            M.apply_f32x64(dev(engine, x), gr)
            ref = oracle.ilu_apply_f32(p, i, dg, f.astype(np.float32), x.astype(np.float32)).astype(np.float64)
            assert np.array_equal(bits(host(gr)), bits(ref)), name
            M.apply_t(dev(engine, x), gr)
            assert np.array_equal(bits(host(gr)), bits(oracle.ilu_apply_t(p, i, dg, f, x))), name
    finally:
        for k, val in zip(keys, saved):
            engine.set_option(k, val)


COL_CASES = {  # name: (matrix, columns, lowest and highest distinct externals of a cell expected)
    # short rows (<= 4 dependencies): loader lanes hold 2 or 4 values each
    "lap3d_40_c16": (lambda: st.laplace3d(40, 40, 40), 16, 1, 128),
    "lap3d_200x200x6_c8": (lambda: st.laplace3d(200, 200, 6), 8, 129, 256),
    # long rows (27-point): 4 or 8 values per loader lane
    "cd27_24_c16": (lambda: st.convdiff27(24, seed=3), 16, 1, 256),
    "cd27_72_c32": (lambda: st.convdiff27(72, seed=3), 32, 257, 512),
}


@pytest.mark.parametrize("name", sorted(COL_CASES))
def test_column_sweep_capacities_bitexact(engine, name):
    """Column-blocked sweeps run (not the level-pool fallback) at every staging capacity, bit-exact."""
    gen, ncol, lo, hi = COL_CASES[name]
    keys = ("tri_col", "tri_col_cols", "tri_col_split")
    saved = [engine.get_option(k) for k in keys]
    try:
        engine.set_option("tri_col", 1)
        engine.set_option("tri_col_cols", ncol)
        engine.set_option("tri_col_split", 1)  # contiguous row blocks: the capacities above are theirs
        p, i, v = gen()
        n = len(p) - 1
        M = H.make_ilu(engine, *upload(engine, p, i, v), want_f32=True)
        for upper in (False, True):
            info = M.sweep_info(upper)
            assert info["column"] and info["ncol"] == ncol, info
            assert lo <= info["max_ext"] <= hi, info
        dg, _ = oracle.diag(p, i)
        f = oracle.ilu0(p, i, v, dg)
        x = st.uniform(23, n)
        gr = engine.vector(n)
        for _ in range(2):
            M.apply(dev(engine, x), gr)
            assert np.array_equal(bits(host(gr)), bits(oracle.ilu_apply(p, i, dg, f, x)))
        M.apply_f32x64(dev(engine, x), gr)
        ref = oracle.ilu_apply_f32(p, i, dg, f.astype(np.float32), x.astype(np.float32)).astype(np.float64)
        assert np.array_equal(bits(host(gr)), bits(ref))
    finally:
        for k, val in zip(keys, saved):
            engine.set_option(k, val)


SPLIT_CASES = {  # name: (matrix, columns, columns per slab of planes (0: automatic), expected split)
    "lap3d_48_c16_auto": (lambda: st.laplace3d(48, 48, 48), 16, 0, 4),
    "lap3d_40x24x32_c64_auto": (lambda: st.laplace3d(40, 24, 32), 64, 0, 8),
    "lap3d_48_c32_s16": (lambda: st.laplace3d(48, 48, 48), 32, 16, 16),
    # 27-point rows reach the next piece of the previous plane: no automatic split
    "cd27_32_c16_auto": (lambda: st.convdiff27(32, seed=5), 16, 0, 1),
    "cd27_32_c16_s4": (lambda: st.convdiff27(32, seed=5), 16, 4, 4),
    "cd27_40_c64_s2": (lambda: st.convdiff27(40, seed=6), 64, 2, 2),
}


@pytest.mark.parametrize("name", sorted(SPLIT_CASES))
def test_column_sweep_plane_split_bitexact(engine, name):
    """Columns cut as pieces of slabs of planes (rows re-ordered inside every level): bit-exact, both precisions."""
    gen, ncol, split, want = SPLIT_CASES[name]
    keys = ("tri_col", "tri_col_cols", "tri_col_split")
    saved = [engine.get_option(k) for k in keys]
    try:
        engine.set_option("tri_col", 1)
        engine.set_option("tri_col_cols", ncol)
        engine.set_option("tri_col_split", split)
        p, i, v = gen()
        n = len(p) - 1
        M = H.make_ilu(engine, *upload(engine, p, i, v), want_f32=True)
        for upper in (False, True):
            info = M.sweep_info(upper)
            assert info["column"] and info["ncol"] == ncol and info["split"] == want, info
        dg, _ = oracle.diag(p, i)
        f = oracle.ilu0(p, i, v, dg)
        x = st.uniform(29, n)
        gr = engine.vector(n)
        for _ in range(2):
            M.apply(dev(engine, x), gr)
            assert np.array_equal(bits(host(gr)), bits(oracle.ilu_apply(p, i, dg, f, x)))
        M.apply_f32x64(dev(engine, x), gr)
        ref = oracle.ilu_apply_f32(p, i, dg, f.astype(np.float32), x.astype(np.float32)).astype(np.float64)
        assert np.array_equal(bits(host(gr)), bits(ref))
        # the symbolic results the API reports keep the (level, row) order
        for upper in (False, True):
            _, lev, _, perm = M.symbolic(upper)
            lev = host(lev)
            assert np.array_equal(host(perm), np.lexsort((np.arange(n), lev)))
    finally:
        for k, val in zip(keys, saved):
            engine.set_option(k, val)


def test_ilu_reference_known_answer(engine):
    # tests/sparse_tests.hpp:423-440
    p = np.array([0, 3, 6, 9], np.int32)
    i = np.array([0, 1, 2, 0, 1, 2, 0, 1, 2], np.int32)
    v = np.array([3.0, 2.0, 1.0, 2.0, 4.0, 3.0, 2.0, 1.0, 6.0])
    M = H.make_ilu(engine, *upload(engine, p, i, v))
    d_, _, _, _ = M.symbolic()
    assert list(host(d_)) == [0, 4, 8]
    ref = np.array([3.0, 2.0, 1.0, 2.0 / 3.0, 8.0 / 3.0, 7.0 / 3.0, 2.0 / 3.0, -0.125, 5.625])
    np.testing.assert_allclose(host(M.get_factor()), ref, rtol=1e-15, atol=0)
    gr = engine.vector(3)
    M.apply(dev(engine, np.array([1.0, 2.0, 3.0])), gr)
    np.testing.assert_allclose(host(gr), [1 / 9, 1 / 9, 4 / 9], rtol=1e-15)


def test_ilu_zero_pivot_reported(engine):
    p = np.array([0, 2, 4], np.int32)
    i = np.array([0, 1, 0, 1], np.int32)
    v = np.array([1.0, 1.0, 1.0, 1.0])  # u_11 = 1 - 1*1 = 0
    with pytest.raises(H.HalaError, match="pivot in row 1"):
        H.make_ilu(engine, *upload(engine, p, i, v))


# ---------------------------------------------------------------- triangular
def test_trsv_reference_cases(engine):
    # tests/sparse_tests.hpp:259-311
    M = 3
    A = np.array([1.0, 4.0, 5.0, 0.0, 2.0, 0.0, 0.0, 0.0, 3.0]).reshape(3, 3, order="F")
    xref = np.array([-3.0, -2.0, -1.0])
    cases = [
        ("L", "N", [0, 1, 3, 5], [0, 0, 1, 0, 2], [1.0, 4.0, 2.0, 5.0, 3.0], "N", A),
        ("L", "T", [0, 1, 3, 5], [0, 0, 1, 0, 2], [1.0, 4.0, 2.0, 5.0, 3.0], "N", A.T),
        ("U", "N", [0, 3, 4, 5], [0, 1, 2, 1, 2], [1.0, 4.0, 5.0, 2.0, 3.0], "N", A.T),
        ("U", "T", [0, 3, 4, 5], [0, 1, 2, 1, 2], [1.0, 4.0, 5.0, 2.0, 3.0], "N", A),
    ]
    for uplo, trans, p, i, v, dg, Aop in cases:
        b = 2.0 * (Aop @ xref)
        gp, gi, gv = upload(engine, np.array(p, np.int32), np.array(i, np.int32), np.array(v))
        T = H.make_triangular_matrix(engine, uplo, dg, gp, gi, gv)
        gx = engine.vector(M)
        T.trsv(trans, 0.5, dev(engine, b), gx)
        np.testing.assert_allclose(host(gx), xref, rtol=1e-14), (uplo, trans)
        ref = oracle.trsv(uplo, trans, dg, 0.5, p, i, v, b)
        assert np.array_equal(bits(host(gx)), bits(ref))


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("dg", ["N", "U"])
def test_trsv_random_triangle(engine, uplo, trans, dg):
    p, i, v = st.laplace3d(12, 12, 12)
    n = len(p) - 1
    rows = []
    for r in range(n):
        c = i[p[r]:p[r + 1]]
        w = v[p[r]:p[r + 1]]
        keep = c <= r if uplo == "L" else c >= r
        rows.append((c[keep], w[keep] * (1.0 + 0.01 * r % 7)))
    tp = np.zeros(n + 1, np.int32)
    tp[1:] = np.cumsum([len(c) for c, _ in rows])
    ti = np.concatenate([c for c, _ in rows]).astype(np.int32)
    tv = np.concatenate([w for _, w in rows])
    b = st.uniform(31, n)
    T = H.make_triangular_matrix(engine, uplo, dg, *upload(engine, tp, ti, tv))
    gx = engine.vector(n)
    for _ in range(2):
        T.trsv(trans, 1.5, dev(engine, b), gx)
        ref = oracle.trsv(uplo, trans, dg, 1.5, tp, ti, tv, b)
        assert np.array_equal(bits(host(gx)), bits(ref))


# ---------------------------------------------------------------- CG
def test_cg_identity_bitexact_history(engine):
    import torch

    p, i, v = st.laplace2d(96, 96)
    n = len(p) - 1
    b = st.uniform(101, n)
    tol = 1e-10 * oracle.cnrm2(b)
    it, x, hist = oracle.cg(p, i, v, b, tol=tol)
    gp, gi, gv = upload(engine, p, i, v)
    gx = engine.vector(n)
    gh = torch.zeros(5000, dtype=torch.float64, device=gx.device)
    git = H.solve_cg(engine, tol, gp, gi, gv, None, dev(engine, b), gx, hist=gh)
    assert git == it
    assert np.array_equal(bits(host(gh)[: it - 1]), bits(hist))
    assert np.array_equal(bits(host(gx)), bits(x))


def test_cg_ilu_published_table(engine):
    # examples/example_gpu.cpp:90-98: CPU ILU-PCG iterations and L2 errors; the GPU path must give
    # the same table and agree with the CPU solution to 1e-11 (:363-365)
    table = {4: (9, 1.000499e-03), 8: (14, 3.092774e-04), 16: (23, 8.669565e-05), 32: (41, 2.300756e-05),
             64: (78, 5.930229e-06), 128: (152, 1.505631e-06)}
    for ns, (iters, err) in table.items():
        p, i, v, f = st.discretize_pde(ns)
        gp, gi, gv = upload(engine, p, i, v)
        ilu = H.make_ilu(engine, gp, gi, gv)
        gu = engine.vector(len(f))
        git = H.solve_cg_ilu(engine, 1e-8, gp, gi, gv, ilu, dev(engine, f), gu)
        u = host(gu)
        assert git == iters
        assert float("%.6e" % st.l2_error(u)) == err
        dg, _ = oracle.diag(p, i)
        it, xo, _ = oracle.cg(p, i, v, f, tol=1e-8, prec=oracle.PREC_ILU, mpntr=p, mindx=i, mdiag=dg,
                              mdata=oracle.ilu0(p, i, v, dg))
        assert np.linalg.norm(u - xo) < 1e-11
        assert np.array_equal(bits(u), bits(xo))


@pytest.mark.parametrize("fuse,xcd,shape", [(1, 1, (32, 32, 32)), (1, 1, (30, 31, 29)), (1, 0, (30, 31, 29)),
                                             (0, 1, (32, 32, 32))],
                         ids=["fused", "fused-ragged", "fused-ragged-dispatch-order", "separate"])
def test_cg_ilu_3d_bitexact(engine, fuse, xcd, shape):
    # fused: the triangular sweeps' gather and scatter passes run inside the
    # CG update and r.z kernels (xcd: the update in XCD-aware block order, 0:
    # dispatch order); separate: the full ILU apply each pass.  Ragged: a row
    # count that is no multiple of the 256-row blocks.
    import torch

    p, i, v = st.laplace3d(*shape)
    n = len(p) - 1
    b = st.uniform(202, n)
    tol = 1e-10 * oracle.cnrm2(b)
    dg, _ = oracle.diag(p, i)
    it, x, hist = oracle.cg(p, i, v, b, tol=tol, prec=oracle.PREC_ILU, mpntr=p, mindx=i, mdiag=dg,
                            mdata=oracle.ilu0(p, i, v, dg))
    gp, gi, gv = upload(engine, p, i, v)
    ilu = H.make_ilu(engine, gp, gi, gv)
    gx = engine.vector(n)
    gh = torch.zeros(2000, dtype=torch.float64, device=gx.device)
    names = ("cg_fuse", "cg_xcd")
    saved = [engine.get_option(k) for k in names]
    for k, val in zip(names, (fuse, xcd)):
        engine.set_option(k, val)
    try:
        git = H.solve_cg_ilu(engine, tol, gp, gi, gv, ilu, dev(engine, b), gx, hist=gh)
    finally:
        for k, val in zip(names, saved):
            engine.set_option(k, val)
    assert git == it
    assert np.array_equal(bits(host(gh)[: it - 1]), bits(hist))
    assert np.array_equal(bits(host(gx)), bits(x))
    r = b - oracle.spmv("N", n, n, 1.0, p, i, v, host(gx), 0.0, np.zeros(n))
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10


def test_cg_jacobi_bitexact(engine):
    p, i, v = st.laplace2d(50, 40, diag=4.3)
    n = len(p) - 1
    d = np.full(n, 4.3)
    b = st.uniform(303, n)
    tol = 1e-10 * oracle.cnrm2(b)
    it, x, _ = oracle.cg(p, i, v, b, tol=tol, prec=oracle.PREC_JACOBI, mdata=d)
    gp, gi, gv = upload(engine, p, i, v)
    gx = engine.vector(n)
    git = H.solve_cg(engine, tol, gp, gi, gv, ("jacobi", dev(engine, d)), dev(engine, b), gx)
    assert git == it
    assert np.array_equal(bits(host(gx)), bits(x))


@pytest.mark.parametrize("prec", ["identity", "jacobi"])
@pytest.mark.parametrize("side", [64, 520])
def test_cg_direction_update_paths(engine, prec, side):
    # up to 2^18 rows the direction update is formed inside the next product,
    # above it by its own pass: both give the oracle's bits (40 iterations)
    import torch

    p, i, v = st.laplace2d(side, side, diag=4.3)
    n = len(p) - 1
    d = np.full(n, 4.3)
    b = st.uniform(404 + side, n)
    kind = oracle.PREC_JACOBI if prec == "jacobi" else oracle.PREC_IDENTITY
    it, x, hist = oracle.cg(p, i, v, b, tol=0.0, max_iter=40, prec=kind, mdata=d)
    gp, gi, gv = upload(engine, p, i, v)
    gx = engine.vector(n)
    gh = torch.zeros(64, dtype=torch.float64, device=gx.device)
    pre = ("jacobi", dev(engine, d)) if prec == "jacobi" else None
    git = H.solve_cg(engine, H.stop_criteria(0.0, 40), gp, gi, gv, pre, dev(engine, b), gx, hist=gh)
    assert git == it == 40
    assert np.array_equal(bits(host(gh)[: it - 1]), bits(hist[: it - 1]))
    assert np.array_equal(bits(host(gx)), bits(x))


def test_cg_grid_stride_product(engine):
    # the fused SpMV + p.Ap launched grid-stride (option spmv_blocks_per_cu > 0) gives the same bits
    import torch

    p, i, v = st.laplace3d(40, 40, 40)
    n = len(p) - 1
    b = st.uniform(505, n)
    tol = 1e-10 * oracle.cnrm2(b)
    gp, gi, gv = upload(engine, p, i, v)
    M = H.make_ilu(engine, gp, gi, gv)
    out = []
    saved = engine.get_option("spmv_blocks_per_cu")
    try:
        for k in (0, 2):
            engine.set_option("spmv_blocks_per_cu", k)
            gx = engine.vector(n)
            gh = torch.zeros(512, dtype=torch.float64, device=gx.device)
            it = H.solve_cg_ilu(engine, tol, gp, gi, gv, M, dev(engine, b), gx, hist=gh)
            out.append((it, bits(host(gx)), bits(host(gh)[: it - 1])))
    finally:
        engine.set_option("spmv_blocks_per_cu", saved)
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
    dg, _ = oracle.diag(p, i)
    f = oracle.ilu0(p, i, v, dg)
    it, x, _ = oracle.cg(p, i, v, b, tol=tol, prec=oracle.PREC_ILU, mpntr=p, mindx=i, mdiag=dg, mdata=f)
    assert out[0][0] == it and np.array_equal(out[0][1], bits(x))


def test_cg_reference_small(engine):
    # tests/solvers_tests.hpp:12-55: tridiagonal -> ILU exact -> two iterations; cyclic 5x5 to 1e-9
    p = np.array([0, 2, 5, 8, 11, 13], np.int32)
    i = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4], np.int32)
    v = np.array([2, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 2], np.float64)
    xref = np.arange(1.0, 6.0)
    A = np.zeros((5, 5))
    for r in range(5):
        A[r, i[p[r]:p[r + 1]]] = v[p[r]:p[r + 1]]
    b = A @ xref
    gp, gi, gv = upload(engine, p, i, v)
    ilu = H.make_ilu(engine, gp, gi, gv)
    gx = engine.vector(5)
    H.solve_cg(engine, H.stop_criteria(0.0, 2), gp, gi, gv, ilu, dev(engine, b), gx)
    assert np.abs(host(gx) - xref).sum() < 1e-14
    p = np.array([0, 3, 6, 9, 12, 15], np.int32)
    i = np.array([0, 1, 4, 0, 1, 2, 1, 2, 3, 2, 3, 4, 0, 3, 4], np.int32)
    v = np.array([2, 1, 1, 1, 2, 1, 1, 2, 1, 1, 2, 1, 1, 1, 2], np.float64)
    A = np.zeros((5, 5))
    for r in range(5):
        A[r, i[p[r]:p[r + 1]]] = v[p[r]:p[r + 1]]
    b = A @ xref
    gp, gi, gv = upload(engine, p, i, v)
    gx = dev(engine, np.array([1.0, 3.0, -4.0, -7.0, 11.0]))
    H.solve_cg_ilu(engine, 1e-9, gp, gi, gv, None, dev(engine, b), gx)
    assert np.abs(host(gx) - xref).sum() / np.linalg.norm(xref) < 1e-9 * 10
