"""CPU tests of matrix ingestion: the oracle's COO-to-CSR step pinned to a
brute-force restatement, and libhala_amd's host Matrix Market reader and
writer (host code, no GPU) on hand-written files of every supported kind."""
import numpy as np
import pytest

import oracle
from hala_amd import mmio
from hala_amd import stencils as st
from hala_amd._lib import HalaError
from tests.conftest import bits


def brute_coo_to_csr(n, m, rows, cols, vals, add_diag):
    """Per row: columns ascending, each column's values summed left to right in input order."""
    acc = [dict() for _ in range(n)]
    for r, c, v in zip(rows.tolist(), cols.tolist(), vals.tolist()):
        acc[r][c] = acc[r][c] + v if c in acc[r] else v
    pntr, indx, out = [0], [], []
    for r in range(n):
        if add_diag and r < m and r not in acc[r]:
            acc[r][r] = 0.0
        for c in sorted(acc[r]):
            indx.append(c)
            out.append(acc[r][c])
        pntr.append(len(indx))
    return np.array(pntr, np.int32), np.array(indx, np.int32), np.array(out, np.float64)


@pytest.mark.parametrize("add_diag", [False, True])
@pytest.mark.parametrize("shape", [(40, 40, 300), (30, 55, 500), (64, 20, 50), (7, 7, 0)])
def test_coo_to_csr_oracle_pins(shape, add_diag):
    n, m, k = shape
    rng = np.random.default_rng(n * 1000 + k)
    rows = rng.integers(0, n, k).astype(np.int32)
    cols = rng.integers(0, m, k).astype(np.int32)
    # duplicates on purpose, values whose sums round (left-to-right order matters)
    rows[: k // 4] = rows[k // 4: 2 * (k // 4)]
    cols[: k // 4] = cols[k // 4: 2 * (k // 4)]
    vals = rng.standard_normal(k) * 10.0 ** rng.integers(-8, 8, k)
    p, i, v = oracle.coo_to_csr(n, m, rows, cols, vals, add_diag)
    bp, bi, bv = brute_coo_to_csr(n, m, rows, cols, vals, add_diag)
    assert np.array_equal(p, bp) and np.array_equal(i, bi)
    assert np.array_equal(bits(v), bits(bv))


def test_coo_to_csr_oracle_rejects_out_of_range():
    with pytest.raises(ValueError, match="entry 2"):
        oracle.coo_to_csr(3, 3, np.array([0, 1, 3]), np.array([0, 1, 1]), np.ones(3))
    with pytest.raises(ValueError, match="entry 0"):
        oracle.coo_to_csr(3, 3, np.array([0]), np.array([-1]), np.ones(1))


def write(tmp_path, name, text):
    f = tmp_path / name
    f.write_text(text)
    return f


def test_mm_coordinate_general(tmp_path):
    f = write(tmp_path, "g.mtx", "%%MatrixMarket matrix coordinate real general\n% comment\n\n3 4 5\n"
              "3 1 -2.5\n1 4 1e-3\n2 2 7\n1 1 4\n3 1 0.5\n")
    h = mmio.info(f)
    assert (h["nrows"], h["ncols"], h["stored"], h["field"], h["symmetry"]) == (3, 4, 5, "real", "general")
    n, m, r, c, v = mmio.read_matrix_market(f)
    assert (n, m) == (3, 4)
    assert r.tolist() == [2, 0, 1, 0, 2] and c.tolist() == [0, 3, 1, 0, 0]
    assert v.tolist() == [-2.5, 1e-3, 7.0, 4.0, 0.5]


def test_mm_symmetric_and_skew_expansion(tmp_path):
    f = write(tmp_path, "s.mtx", "%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n1 1 2\n2 1 -1\n3 3 5\n3 2 4\n")
    _, _, r, c, v = mmio.read_matrix_market(f)
    assert list(zip(r.tolist(), c.tolist(), v.tolist())) == [(0, 0, 2.0), (1, 0, -1.0), (0, 1, -1.0), (2, 2, 5.0),
                                                              (2, 1, 4.0), (1, 2, 4.0)]
    _, _, r, c, v = mmio.read_matrix_market(f, expand=False)
    assert r.size == 4
    f = write(tmp_path, "k.mtx", "%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 3\n3 1 -1.5\n")
    _, _, r, c, v = mmio.read_matrix_market(f)
    assert list(zip(r.tolist(), c.tolist(), v.tolist())) == [(1, 0, 3.0), (0, 1, -3.0), (2, 0, -1.5), (0, 2, 1.5)]


def test_mm_pattern_integer_array(tmp_path):
    f = write(tmp_path, "p.mtx", "%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 2\n2 1\n")
    _, _, r, c, v = mmio.read_matrix_market(f)
    assert r.tolist() == [0, 1] and c.tolist() == [1, 0] and v.tolist() == [1.0, 1.0]
    f = write(tmp_path, "i.mtx", "%%MatrixMarket matrix coordinate integer general\n2 2 1\n2 2 -7\n")
    assert mmio.read_matrix_market(f)[4].tolist() == [-7.0]
    # array: column-major, every position
    f = write(tmp_path, "a.mtx", "%%MatrixMarket matrix array real general\n2 3\n1\n2\n3\n4\n5\n6\n")
    n, m, r, c, v = mmio.read_matrix_market(f)
    assert (n, m) == (2, 3) and r.tolist() == [0, 1, 0, 1, 0, 1] and c.tolist() == [0, 0, 1, 1, 2, 2]
    assert v.tolist() == [1, 2, 3, 4, 5, 6]
    # symmetric array: lower triangle column by column
    f = write(tmp_path, "as.mtx", "%%MatrixMarket matrix array real symmetric\n3 3\n1\n2\n3\n4\n5\n6\n")
    _, _, r, c, v = mmio.read_matrix_market(f)
    got = {(a, b): x for a, b, x in zip(r.tolist(), c.tolist(), v.tolist())}
    A = np.array([[1, 2, 3], [2, 4, 5], [3, 5, 6]], float)
    assert len(got) == 9 and all(got[(a, b)] == A[a, b] for a in range(3) for b in range(3))


@pytest.mark.parametrize("text, msg", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", "field"),
    ("%%MatrixMarket matrix coordinate real hermitian\n1 1 1\n1 1 1\n", "symmetry"),
    ("%%MatrixMarket vector coordinate real general\n1 1 1\n1 1 1\n", "object"),
    ("%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n", "not a Matrix Market"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n", "ends after 1"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", "out of range"),
    ("%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 1 1\n", "not square"),
])
def test_mm_rejects(tmp_path, text, msg):
    f = write(tmp_path, "bad.mtx", text)
    with pytest.raises(HalaError, match=msg):
        mmio.read_matrix_market(f)


def test_mm_write_read_roundtrip(tmp_path):
    p, i, v = st.convdiff27(6, seed=4)
    f = tmp_path / "c.mtx"
    mmio.write_matrix_market(f, p, i, v)
    n, m, r, c, w = mmio.read_matrix_market(f)
    p2, i2, v2 = oracle.coo_to_csr(n, m, r, c, w)
    assert np.array_equal(p2, p) and np.array_equal(i2, i) and np.array_equal(bits(v2), bits(v))
