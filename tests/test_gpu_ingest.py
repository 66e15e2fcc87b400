"""GPU tests of matrix ingestion: hala_coo_to_csr bit-exact against the
oracle, its range check, and Matrix Market files loaded to device CSR and
solved (the canonical CSR of a shuffled, duplicated file is the original one)."""
import numpy as np
import pytest

import hala_amd as H
import oracle
from hala_amd import mmio
from hala_amd import stencils as st
from hala_amd._lib import HalaError
from tests.conftest import bits

pytestmark = pytest.mark.gpu


def host(t):
    return t.cpu().numpy()


CASES = {  # name: (nrows, ncols, entries, seed)
    "square_dups": (300, 300, 4000, 1),
    "wide": (120, 500, 3000, 2),
    "tall_sparse": (700, 90, 400, 3),
    "long_rows": (16, 16, 5000, 4),
    "empty": (50, 50, 0, 5),
}


@pytest.mark.parametrize("add_diag", [False, True])
@pytest.mark.parametrize("name", sorted(CASES))
def test_coo_to_csr_bitexact(engine, name, add_diag):
    n, m, k, seed = CASES[name]
    print(f"[coo_to_csr] {name} add_diag={add_diag}", flush=True)
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n, k).astype(np.int32)
    cols = rng.integers(0, m, k).astype(np.int32)
    vals = rng.standard_normal(k) * 10.0 ** rng.integers(-6, 6, k)
    p, i, v = mmio.coo_to_csr(engine, n, m, rows, cols, vals, add_diag)
    rp, ri, rv = oracle.coo_to_csr(n, m, rows, cols, vals, add_diag)
    assert np.array_equal(host(p), rp) and np.array_equal(host(i), ri)
    assert np.array_equal(bits(host(v)), bits(rv))


def test_coo_to_csr_rejects_out_of_range(engine):
    with pytest.raises(HalaError, match="entry 3") as e:
        mmio.coo_to_csr(engine, 4, 4, [0, 1, 2, 2], [0, 1, 2, 4], [1.0, 1.0, 1.0, 1.0])
    assert e.value.code == -4


def test_matrix_market_shuffled_file_solves(engine, tmp_path):
    """A 3D Laplacian written with shuffled, split (duplicated) entries loads back to its own CSR and the
    ILU-PCG solve through it equals the solve of the original arrays bit for bit."""
    p, i, v = st.laplace3d(20, 20, 20)
    n = p.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(p))
    rng = np.random.default_rng(9)
    # split every value into a + (v - a) with a exact (a multiple of 1/4): sums restore v exactly
    a = np.round(v * 4.0) / 4.0 - 0.25
    r2, c2, v2 = np.concatenate([rows, rows]), np.concatenate([i, i]), np.concatenate([a, v - a])
    order = rng.permutation(r2.size)
    f = tmp_path / "lap.mtx"
    with open(f, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate real general\n% shuffled, split entries\n")
        fh.write(f"{n} {n} {r2.size}\n")
        for q in order:
            fh.write(f"{r2[q] + 1} {c2[q] + 1} {float(v2[q])!r}\n")
    nr, nc, gp, gi, gv = mmio.load_matrix_market(engine, f)
    assert (nr, nc) == (n, n)
    assert np.array_equal(host(gp), p) and np.array_equal(host(gi), i) and np.array_equal(bits(host(gv)), bits(v))
    b = st.uniform(7, n)
    import torch

    outs = []
    for P, I, V in ((gp, gi, gv), tuple(torch.as_tensor(x).cuda() for x in (p, i, v))):
        x = engine.vector(n)
        it = H.solve_cg_ilu(engine, 1e-10, P, I, V, None, torch.as_tensor(b).cuda(), x)
        outs.append((it, host(x)))
    assert outs[0][0] == outs[1][0] and np.array_equal(bits(outs[0][1]), bits(outs[1][1]))


def test_missing_diagonal_rejected_then_inserted(engine, tmp_path):
    """ILU(0) needs every diagonal entry (sparse/hala_sparse_core.hpp:42): a pattern without one is
    refused with HALA_ERR_PATTERN; add_diag inserts explicit zeros and the factorisation then runs
    (row 1's pivot comes from the elimination: 0 - (-1/4)(-1) = -1/4)."""
    f = tmp_path / "nodiag.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real symmetric\n4 4 6\n1 1 4\n2 1 -1\n3 2 -1\n3 3 4\n"
                 "4 3 -1\n4 4 4\n")
    n, m, gp, gi, gv = mmio.load_matrix_market(engine, f)
    with pytest.raises(HalaError) as e:
        H.make_ilu(engine, gp, gi, gv)
    assert e.value.code == -4
    n, m, gp, gi, gv = mmio.load_matrix_market(engine, f, add_diag=True)
    assert host(gp).tolist() == [0, 2, 5, 8, 10]
    assert host(gi).tolist() == [0, 1, 0, 1, 2, 1, 2, 3, 2, 3]
    assert host(gv)[3] == 0.0
    M = H.make_ilu(engine, gp, gi, gv)
    dg, _ = oracle.diag(host(gp), host(gi))
    f = oracle.ilu0(host(gp), host(gi), host(gv), dg)
    assert f[3] == -0.25 and np.array_equal(bits(host(M.get_factor())), bits(f))
