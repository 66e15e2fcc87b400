"""Matrix ingestion: Matrix Market files and COO triplets to the sorted CSR the
solvers take (ascending unique columns, optionally every diagonal present).

read_matrix_market / write_matrix_market work on host arrays (libhala_amd's
host reader and writer); coo_to_csr runs on the device (hala_coo_to_csr);
load_matrix_market chains the two.  The reference has no ingestion: its CSR
arguments must already be canonical (sparse/hala_sparse_core.hpp:42, :56-59).
"""
import ctypes as C

import numpy as np

from ._lib import call

FIELDS = {0: "real", 1: "integer", 2: "pattern"}
SYMMETRIES = {0: "general", 1: "symmetric", 2: "skew-symmetric"}


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def info(path):
    """Header of a Matrix Market file: dict(nrows, ncols, stored, field, symmetry)."""
    r, c, z, e, f, s = C.c_int(), C.c_int(), C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
    call("hala_mm_info", str(path).encode(), C.byref(r), C.byref(c), C.byref(z), C.byref(e), C.byref(f), C.byref(s))
    return {"nrows": r.value, "ncols": c.value, "stored": z.value, "bound": e.value, "field": FIELDS[f.value],
            "symmetry": SYMMETRIES[s.value]}


def read_matrix_market(path, expand=True):
    """(nrows, ncols, rows, cols, vals): 0-based COO host arrays in file order; symmetric and
    skew-symmetric files expanded (mirror after its entry) unless expand is False."""
    h = info(path)
    cap = max(h["bound"], 1)
    rows, cols = np.empty(cap, np.int32), np.empty(cap, np.int32)
    vals = np.empty(cap, np.float64)
    cnt = C.c_int64()
    call("hala_mm_read_coo", str(path).encode(), int(bool(expand)), cap, _p(rows), _p(cols), _p(vals), C.byref(cnt))
    k = cnt.value
    return h["nrows"], h["ncols"], rows[:k].copy(), cols[:k].copy(), vals[:k].copy()


def write_matrix_market(path, pntr, indx, vals, ncols=None):
    """Host CSR to a 'coordinate real general' file (17 significant digits: values round-trip)."""
    pntr = np.ascontiguousarray(pntr, np.int32)
    indx = np.ascontiguousarray(indx, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    n = pntr.size - 1
    call("hala_mm_write_csr", str(path).encode(), n, n if ncols is None else int(ncols), _p(pntr), _p(indx),
         _p(vals))


def coo_to_csr(engine, nrows, ncols, rows, cols, vals, add_diag=False):
    """Device CSR (pntr, indx, vals tensors on the engine's device) of COO triplets (tensors or
    arrays): (row, column, position) order, duplicates summed left to right in input order,
    add_diag inserts a zero for every missing diagonal entry."""
    import torch

    from . import _ptr

    dev = torch.device("cuda", engine.device())
    r = torch.as_tensor(rows, dtype=torch.int32).to(dev).contiguous()
    c = torch.as_tensor(cols, dtype=torch.int32).to(dev).contiguous()
    v = torch.as_tensor(vals, dtype=torch.float64).to(dev).contiguous()
    pntr = torch.empty(nrows + 1, dtype=torch.int32, device=dev)
    nnz = C.c_int64()
    args = (engine.handle, int(nrows), int(ncols), r.numel(), _ptr(r), _ptr(c), _ptr(v), int(bool(add_diag)))
    call("hala_coo_to_csr", *args, _ptr(pntr), None, None, C.byref(nnz))
    indx = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)[: nnz.value]
    out = torch.empty(max(nnz.value, 1), dtype=torch.float64, device=dev)[: nnz.value]
    call("hala_coo_to_csr", *args, _ptr(pntr), _ptr(indx), _ptr(out), C.byref(nnz))
    return pntr, indx, out


def load_matrix_market(engine, path, add_diag=False):
    """Matrix Market file to device CSR (see coo_to_csr); returns (nrows, ncols, pntr, indx, vals)."""
    n, m, rows, cols, vals = read_matrix_market(path)
    p, i, v = coo_to_csr(engine, n, m, rows, cols, vals, add_diag)
    return n, m, p, i, v
