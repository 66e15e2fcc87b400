"""hala_amd -- HALA's sparse iterative-solve path on MI355X (gfx950).

Python mirror of the reference's engine/matrix/ILU/solver interface
(gpu/hala_gpu_engine.hpp, gpu/hala_rocm_sparse_general.hpp,
gpu/hala_gpu_ilu.hpp, hex/solvers/*.hpp), used by the tests, the benchmark
and the autograd harness.  Device memory comes from torch tensors; every
computation is a call into libhala_amd.so through its C ABI.  The C++ drop-in
for user code calling hala:: is include/hala.hpp and include/hala_solvers.hpp.
"""
import ctypes as C

import numpy as np

from ._lib import CsrInfo, HalaError, IluInfo, call, lib  # noqa: F401
from . import stencils  # noqa: F401

__all__ = [
    "HalaError", "gpu_engine", "stop_criteria", "make_sparse_matrix", "sparse_gemv", "make_ilu",
    "make_triangular_matrix", "solve_cg", "solve_cg_ilu", "solve_gmres", "solve_gmres_ilu",
    "solve_gmres_ir", "solve_batch_cg", "solve_batch_cg_ilu", "dot", "norm2", "axpy", "scal", "vcopy", "stencils",
]


def _torch():
    import torch
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class stop_criteria:
    """hex/solvers/hala_solvers_core.hpp:52-71: absolute tolerance and/or iteration cap."""

    def __init__(self, tol=0.0, max_iter=2**31 - 1):
        self.tol = float(tol)
        self.max_iter = int(max_iter)


class gpu_engine:
    """gpu/hala_gpu_engine.hpp:167-307: one device, one stream (torch's current stream by default)."""

    def __init__(self, device=0, stream=None):
        torch = _torch()
        self._device = int(device)
        if stream is None:
            stream = torch.cuda.current_stream(self._device)
        self.stream = stream
        h = C.c_void_p()
        call("hala_ctx_create", self._device, C.c_void_p(stream.cuda_stream), C.byref(h))
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None) is not None and callable(lib):
            lib().hala_ctx_destroy(self.handle)
            self.handle = None

    def device(self):
        return self._device

    def set_option(self, name, value):
        call("hala_ctx_set_option", self.handle, name.encode(), int(value))

    def get_option(self, name):
        v = C.c_int()
        call("hala_ctx_get_option", self.handle, name.encode(), C.byref(v))
        return v.value

    def synchronize(self):
        call("hala_ctx_sync", self.handle)

    def load(self, arr, dtype=None):
        torch = _torch()
        a = np.ascontiguousarray(arr if dtype is None else np.asarray(arr, dtype=dtype))
        return torch.from_numpy(a).to(f"cuda:{self._device}")

    @staticmethod
    def unload(t):
        return t.detach().cpu().numpy()

    def vector(self, n, value=0.0, dtype=None):
        torch = _torch()
        dtype = dtype or torch.float64
        t = torch.empty(int(n), dtype=dtype, device=f"cuda:{self._device}")
        if n:
            if dtype == torch.float64:
                call("hala_fill_f64", self.handle, _ptr(t), int(n), float(value))
            elif dtype == torch.float32:
                call("hala_fill_f32", self.handle, _ptr(t), int(n), float(value))
            else:
                call("hala_fill_i32", self.handle, _ptr(t), int(n), int(value))
        return t


def _dt(t):
    torch = _torch()
    return {torch.float64: 64, torch.float32: 32}[t.dtype]


class gpu_sparse_matrix:
    """gpu/hala_rocm_sparse_general.hpp:74-169: non-owning view of a device CSR (kept alive here)."""

    def __init__(self, engine, nrows, ncols, pntr, indx, vals):
        self.engine = engine
        self.pntr, self.indx, self.vals = pntr, indx, vals
        self.nrows, self.ncols = int(nrows), int(ncols)
        h = C.c_void_p()
        call("hala_csr_create", engine.handle, self.nrows, self.ncols, int(indx.numel()), _ptr(pntr), _ptr(indx),
             C.byref(h))
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None) is not None and callable(lib):
            lib().hala_csr_destroy(self.handle)
            self.handle = None

    def info(self):
        inf = CsrInfo()
        call("hala_csr_info", self.handle, C.byref(inf))
        return {k: getattr(inf, k) for k, _ in CsrInfo._fields_}

    def set_kernel(self, k):
        call("hala_csr_set_kernel", self.handle, int(k))

    def gemv(self, trans, alpha, x, beta, y, vals=None):
        vals = self.vals if vals is None else vals
        t = trans.encode() if isinstance(trans, str) else trans
        if _dt(vals) == 64:
            call("hala_spmv_f64", self.handle, t, float(alpha), _ptr(vals), _ptr(x), float(beta), _ptr(y))
        elif _dt(x) == 32:
            call("hala_spmv_f32", self.handle, t, float(alpha), _ptr(vals), _ptr(x), float(beta), _ptr(y))
        else:
            assert t in (b"N", b"n"), "FP32-value product with FP64 vectors is 'N' only"
            call("hala_spmv_f32x64", self.handle, float(alpha), _ptr(vals), _ptr(x), float(beta), _ptr(y))
        return y

    def gemm(self, transa, transb, b_rows, b_cols, alpha, B, ldb, beta, C, ldc):
        """C = alpha op(A) op(B) + beta C (column-major B, C; gpu_sparse_matrix::gemm)."""
        import torch

        sf = {torch.float64: "f64", torch.float32: "f32", torch.complex128: "z",
              torch.complex64: "c"}[self.vals.dtype]
        if sf in ("z", "c"):
            a, b = complex(alpha), complex(beta)
            sc = (a.real, a.imag), (b.real, b.imag)
        else:
            sc = (float(alpha),), (float(beta),)
        call("hala_spmm_" + sf, self.handle, transa.encode(), transb.encode(), int(b_rows), int(b_cols), *sc[0],
             _ptr(self.vals), _ptr(B), int(ldb), *sc[1], _ptr(C), int(ldc))
        return C


def make_sparse_matrix(engine, ncols, pntr, indx, vals):
    """make_sparse_matrix(gpu_engine, ncols, pntr, indx, vals): gpu/hala_rocm_sparse_general.hpp:179-185."""
    return gpu_sparse_matrix(engine, int(pntr.numel()) - 1, ncols, pntr, indx, vals)


def sparse_gemv(engine, trans, M, N, alpha, pntr, indx, vals, x, beta, y):
    """sparse_gemv(gpu_engine, ...): gpu/hala_rocm_sparse_general.hpp:188-199."""
    return gpu_sparse_matrix(engine, M, N, pntr, indx, vals).gemv(trans, alpha, x, beta, y)


class gpu_ilu:
    """gpu/hala_gpu_ilu.hpp:45-185: ILU(0) factor owned on the device, pattern aliased."""

    def __init__(self, engine, pntr, indx, vals, policy="N", want_f32=False):
        self.engine = engine
        self.pntr, self.indx = pntr, indx
        self.n = int(pntr.numel()) - 1
        self.nnz = int(indx.numel())
        h = C.c_void_p()
        call("hala_ilu_create", engine.handle, self.n, self.nnz, _ptr(pntr), _ptr(indx), C.byref(h))
        self.handle = h
        self.policy = policy
        if vals is not None:
            self.factor(vals, want_f32)

    def __del__(self):
        if getattr(self, "handle", None) is not None and callable(lib):
            lib().hala_ilu_destroy(self.handle)
            self.handle = None

    def factor(self, vals, want_f32=False):
        call("hala_ilu_factor_f64", self.handle, _ptr(vals), int(bool(want_f32)))

    def info(self):
        inf = IluInfo()
        call("hala_ilu_info", self.handle, C.byref(inf))
        return {k: getattr(inf, k) for k, _ in IluInfo._fields_}

    def sweep_info(self, upper=False):
        """Layout of one triangular sweep: column-blocked or level pool, and its shape."""
        o = (C.c_int * 8)()
        call("hala_ilu_sweep_info", self.handle, 1 if upper else 0, o)
        return {"column": bool(o[0]), "ncol": o[1], "cells": o[2], "window": o[3], "max_ext": o[4], "why": o[5],
                "split": o[6], "plane": o[7]}

    def stencil_info(self):
        """Pencil schedule of a stencil pattern: kind (0 none, 1 5-pt, 2 9-pt, 3 7-pt, 4 27-pt), grid, in use."""
        o = (C.c_int * 5)()
        call("hala_ilu_stencil_info", self.handle, o)
        return {"kind": o[0], "nx": o[1], "ny": o[2], "nz": o[3], "in_use": bool(o[4])}

    def apply(self, x, r, num_rhs=1):
        assert num_rhs == 1
        if _dt(x) == 64:
            call("hala_ilu_apply_f64", self.handle, _ptr(x), _ptr(r))
        else:
            call("hala_ilu_apply_f32", self.handle, _ptr(x), _ptr(r))
        return r

    def apply_f32x64(self, x, r):
        call("hala_ilu_apply_f32x64", self.handle, _ptr(x), _ptr(r))
        return r

    def apply_t(self, x, r):
        call("hala_ilu_apply_t_f64", self.handle, _ptr(x), _ptr(r))
        return r

    def symbolic(self, upper=False):
        torch = _torch()
        dev = self.pntr.device
        inf = self.info()
        nlev = inf["nlev_upper" if upper else "nlev_lower"]
        diag = torch.empty(self.n, dtype=torch.int32, device=dev)
        lev = torch.empty(self.n, dtype=torch.int32, device=dev)
        lptr = torch.empty(nlev + 1, dtype=torch.int32, device=dev)
        perm = torch.empty(self.n, dtype=torch.int32, device=dev)
        call("hala_ilu_get_symbolic", self.handle, int(bool(upper)), _ptr(diag), _ptr(lev), _ptr(lptr), _ptr(perm))
        return diag, lev, lptr, perm

    def get_factor(self):
        torch = _torch()
        f = torch.empty(self.nnz, dtype=torch.float64, device=self.pntr.device)
        call("hala_ilu_get_factor_f64", self.handle, _ptr(f))
        return f


def make_ilu(engine, pntr, indx, vals, policy="N", want_f32=False):
    """make_ilu(gpu_engine, pntr, indx, vals, policy): gpu/hala_gpu_ilu.hpp:191-195."""
    return gpu_ilu(engine, pntr, indx, vals, policy, want_f32)


class gpu_triangular_matrix:
    """gpu/hala_rocm_sparse_triangular.hpp:29-178 with sparse_trsv_array semantics."""

    def __init__(self, engine, uplo, diag, pntr, indx, vals):
        self.engine = engine
        self.pntr, self.indx, self.vals = pntr, indx, vals
        self.n = int(pntr.numel()) - 1
        h = C.c_void_p()
        call("hala_tri_create", engine.handle, uplo.encode(), diag.encode(), self.n, int(indx.numel()), _ptr(pntr),
             _ptr(indx), C.byref(h))
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None) is not None and callable(lib):
            lib().hala_tri_destroy(self.handle)
            self.handle = None

    def _suffix(self):
        import torch

        return {torch.float64: "f64", torch.float32: "f32", torch.complex128: "z", torch.complex64: "c"}[self.vals.dtype]

    def _alpha(self, sf, alpha):
        if sf in ("z", "c"):
            a = complex(alpha)
            return (a.real, a.imag)
        return (float(alpha),)

    def trsv(self, trans, alpha, b, x):
        """x = op(T)^-1 (alpha b); the scalar type follows the matrix values."""
        sf = self._suffix()
        call("hala_trsv_" + sf, self.handle, trans.encode(), *self._alpha(sf, alpha), _ptr(self.vals), _ptr(b),
             _ptr(x))
        return x

    def trsm(self, transa, transb, nrhs, alpha, B, ldb=-1):
        """B <- op(T)^-1 (alpha B), column by column (sparse_trsm)."""
        sf = self._suffix()
        call("hala_trsm_" + sf, self.handle, transa.encode(), transb.encode(), int(nrhs), *self._alpha(sf, alpha),
             _ptr(self.vals), _ptr(B), int(ldb))
        return B


def make_triangular_matrix(engine, uplo, diag, pntr, indx, vals):
    return gpu_triangular_matrix(engine, uplo, diag, pntr, indx, vals)


# ---- BLAS-1 on the engine (gpu/hala_gpu_blas1.hpp) ---------------------------
def dot(engine, x, y):
    if _dt(x) == 64:
        r = C.c_double()
        call("hala_dot_f64", engine.handle, int(x.numel()), _ptr(x), _ptr(y), C.byref(r))
    else:
        r = C.c_float()
        call("hala_dot_f32", engine.handle, int(x.numel()), _ptr(x), _ptr(y), C.byref(r))
    return r.value


def norm2(engine, x):
    if _dt(x) == 64:
        r = C.c_double()
        call("hala_nrm2_f64", engine.handle, int(x.numel()), _ptr(x), C.byref(r))
    else:
        r = C.c_float()
        call("hala_nrm2_f32", engine.handle, int(x.numel()), _ptr(x), C.byref(r))
    return r.value


def axpy(engine, alpha, x, y):
    f = "hala_axpy_f64" if _dt(x) == 64 else "hala_axpy_f32"
    call(f, engine.handle, int(x.numel()), float(alpha), _ptr(x), _ptr(y))
    return y


def scal(engine, alpha, x):
    f = "hala_scal_f64" if _dt(x) == 64 else "hala_scal_f32"
    call(f, engine.handle, int(x.numel()), float(alpha), _ptr(x))
    return x


def vcopy(engine, x, y):
    f = "hala_copy_f64" if _dt(x) == 64 else "hala_copy_f32"
    call(f, engine.handle, int(x.numel()), _ptr(x), _ptr(y))
    return y


# ---- solvers (hex/solvers) ---------------------------------------------------
def _stop(stop):
    return stop if isinstance(stop, stop_criteria) else stop_criteria(stop)


def solve_cg(engine, stop, pntr, indx, vals, precon, b, x, hist=None, matrix=None):
    """solve_cg(engine, stop, pntr, indx, vals, precon, b, x): hex/solvers/hala_solvers_cg.hpp:230-244.

    precon: None (identity copy), a gpu_ilu, or ("jacobi", diagonal_tensor).
    Returns the iteration count (operator applications, counted from 1).
    """
    stop = _stop(stop)
    A = matrix or gpu_sparse_matrix(engine, int(pntr.numel()) - 1, int(pntr.numel()) - 1, pntr, indx, vals)
    it = C.c_int()
    if precon is None:
        kind, ilu, dg = 0, None, None
    elif isinstance(precon, gpu_ilu):
        kind, ilu, dg = 1, precon, None
    else:
        kind, ilu, dg = 2, None, precon[1]
    nh = int(hist.numel()) if hist is not None else 0
    call("hala_cg_f64", A.handle, _ptr(vals), kind, ilu.handle if ilu else C.c_void_p(0), _ptr(dg), _ptr(b),
         _ptr(x), stop.tol, stop.max_iter, C.byref(it), _ptr(hist), nh)
    return it.value


def solve_cg_ilu(engine, stop, pntr, indx, vals, ilu, b, x, hist=None, matrix=None):
    """solve_cg_ilu(engine, stop, pntr, indx, vals, ilu, b, x): hex/solvers/hala_solvers_cg.hpp:312-326."""
    if ilu is None:
        ilu = make_ilu(engine, pntr, indx, vals)
    return solve_cg(engine, stop, pntr, indx, vals, ilu, b, x, hist=hist, matrix=matrix)


def solve_batch_cg(engine, stop, pntr, indx, vals, precon, B, X, hist=None, matrix=None):
    """solve_batch_cg(engine, stop, pntr, indx, vals, precon, B, X): hex/solvers/hala_solvers_cg_batch.hpp:68-110.

    B, X: n*nrhs entries, right-hand side k at [k n, (k+1) n) (a (nrhs, n) tensor works);
    precon: None (identity copy) or a gpu_ilu.  Returns the iteration count.
    """
    stop = _stop(stop)
    n = int(pntr.numel()) - 1
    assert B.numel() % n == 0 and X.numel() == B.numel(), "B and X must hold n * nrhs entries"
    A = matrix or gpu_sparse_matrix(engine, n, n, pntr, indx, vals)
    it = C.c_int()
    nh = int(hist.numel()) if hist is not None else 0
    call("hala_cg_batch_f64", A.handle, _ptr(vals), 0 if precon is None else 1,
         precon.handle if precon is not None else C.c_void_p(0), int(B.numel()) // n, _ptr(B), _ptr(X), stop.tol,
         stop.max_iter, C.byref(it), _ptr(hist), nh)
    return it.value


def solve_batch_cg_ilu(engine, stop, pntr, indx, vals, ilu, B, X, hist=None, matrix=None):
    """solve_batch_cg_ilu: hex/solvers/hala_solvers_cg_batch.hpp:156-249 (ILU built on the fly when None)."""
    if ilu is None:
        ilu = make_ilu(engine, pntr, indx, vals)
    return solve_batch_cg(engine, stop, pntr, indx, vals, ilu, B, X, hist=hist, matrix=matrix)


def solve_cg_t(engine, stop, matrix, vals, ilu, b, x, hist=None):
    """Adjoint solve A^T x = b preconditioned with (LU)^-T."""
    stop = _stop(stop)
    it = C.c_int()
    nh = int(hist.numel()) if hist is not None else 0
    call("hala_cg_t_f64", matrix.handle, _ptr(vals), ilu.handle, _ptr(b), _ptr(x), stop.tol, stop.max_iter,
         C.byref(it), _ptr(hist), nh)
    return it.value


def solve_gmres(engine, stop, restart, pntr, indx, vals, precon, b, x, hist=None, matrix=None):
    """solve_gmres(engine, stop, restart, ...): hex/solvers/hala_solvers_gmres.hpp:125-230.

    precon: None (identity) or a gpu_ilu.  stop.max_iter counts restart cycles.
    Returns (preconditioner applications, number of Givens estimates recorded).
    """
    stop = _stop(stop)
    A = matrix or gpu_sparse_matrix(engine, int(pntr.numel()) - 1, int(pntr.numel()) - 1, pntr, indx, vals)
    tot, nh = C.c_int(), C.c_int()
    kind = 0 if precon is None else 1
    mh = int(hist.numel()) if hist is not None else 0
    call("hala_gmres_f64", A.handle, _ptr(vals), kind, precon.handle if precon is not None else C.c_void_p(0),
         _ptr(b), _ptr(x), stop.tol, int(restart), stop.max_iter, C.byref(tot), _ptr(hist), mh, C.byref(nh))
    return tot.value, nh.value


def solve_gmres_ilu(engine, stop, restart, pntr, indx, vals, ilu, b, x, hist=None, matrix=None):
    """solve_gmres_ilu: hex/solvers/hala_solvers_gmres.hpp:250-267."""
    if ilu is None:
        ilu = make_ilu(engine, pntr, indx, vals)
    return solve_gmres(engine, stop, restart, pntr, indx, vals, ilu, b, x, hist=hist, matrix=matrix)


def solve_gmres_ir(engine, matrix, vals, vals32, ilu, b, x, rtol=1e-10, inner_rtol=1e-4, restart=30,
                   max_refine=50, hist=None):
    """FP32 ILU(0) + FP32 product inside GMRES(restart), FP64 refinement to ||b-Ax|| < rtol ||b||.

    Returns (inner preconditioner applications, outer residuals recorded).
    """
    tot, nh = C.c_int(), C.c_int()
    mh = int(hist.numel()) if hist is not None else 0
    call("hala_gmres_ir_f64", matrix.handle, _ptr(vals), _ptr(vals32), ilu.handle, _ptr(b), _ptr(x), float(rtol),
         float(inner_rtol), int(restart), int(max_refine), C.byref(tot), _ptr(hist), mh, C.byref(nh))
    return tot.value, nh.value


# ---- batched independent systems -------------------------------------------------
class batch_solver:
    """Jacobi-PCG over many independent packed CSR systems in one persistent launch."""

    def __init__(self, engine, row_off, pntr, indx):
        self.engine = engine
        self.row_off, self.pntr, self.indx = row_off, pntr, indx
        self.nsys = int(row_off.numel()) - 1
        h = C.c_void_p()
        call("hala_batch_create", engine.handle, self.nsys, _ptr(row_off), _ptr(pntr), _ptr(indx), C.byref(h))
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None) is not None and callable(lib):
            lib().hala_batch_destroy(self.handle)
            self.handle = None

    def set_shape(self, team, wg_per_cu):
        call("hala_batch_set_shape", self.handle, int(team), int(wg_per_cu))

    def shape(self):
        t, w = C.c_int(), C.c_int()
        call("hala_batch_get_shape", self.handle, C.byref(t), C.byref(w))
        return t.value, w.value

    def solve(self, vals, dvals, b, x, rtol=1e-10, max_iter=100000, iters=None):
        call("hala_batch_pcg_f64", self.handle, _ptr(vals), _ptr(dvals), _ptr(b), _ptr(x), float(rtol), int(max_iter),
             _ptr(iters))
        return x

    def tune(self, vals, dvals, b, x, rtol=1e-10, max_iter=100000, reps=2):
        t, w, ms = C.c_int(), C.c_int(), C.c_double()
        call("hala_batch_tune", self.handle, _ptr(vals), _ptr(dvals), _ptr(b), _ptr(x), float(rtol), int(max_iter),
             int(reps), C.byref(t), C.byref(w), C.byref(ms))
        return t.value, w.value, ms.value


# ---- adjoint pieces ------------------------------------------------------------
def grad_vals(matrix, lam, x, gvals):
    """gvals[j] = -lam[row(j)] * x[col(j)]."""
    call("hala_grad_vals_f64", matrix.handle, _ptr(lam), _ptr(x), _ptr(gvals))
    return gvals


def grad_coef(engine, cptr, cnnz, ccoef, gvals, gcoef):
    """gcoef[c] = sum_e ccoef[e] * gvals[cnnz[e]] over e in cptr[c]..cptr[c+1]."""
    call("hala_grad_coef_f64", engine.handle, int(cptr.numel()) - 1, _ptr(cptr), _ptr(cnnz), _ptr(ccoef), _ptr(gvals),
         _ptr(gcoef))
    return gcoef


def assemble(engine, aptr, aidx, aw, coef, vals):
    """vals[j] = sum_e aw[e] * coef[aidx[e]] over e in aptr[j]..aptr[j+1]."""
    call("hala_assemble_f64", engine.handle, int(aptr.numel()) - 1, _ptr(aptr), _ptr(aidx), _ptr(aw), _ptr(coef),
         _ptr(vals))
    return vals


# ---- mixed engine (wax/hala_lib_extensions.hpp:68-376, wax/hala_mixed_overloads.hpp) ----
class mixed_engine:
    """Host (numpy) inputs and outputs, the work on the GPU: every call loads its numpy arguments
    into device tensors, runs the gpu_engine path and copies its outputs back in place."""

    def __init__(self, gpu):
        self._gpu = gpu

    def gpu(self):
        return self._gpu

    def device(self):
        return self._gpu.device()


class mixed_ilu:
    """ILU(0) of host arrays (wax/hala_lib_extensions.hpp:317-376): the device factor keeps its own
    copies of the pattern; apply() takes and fills numpy arrays."""

    def __init__(self, engine, gilu):
        self.engine, self._gilu = engine, gilu

    def cilu(self):
        return self._gilu

    def apply(self, x, r, num_rhs=1):
        gx = _to_dev(self.engine, x)
        gr = self.engine.gpu().vector(gx.numel())
        n = self._gilu.n
        for k in range(num_rhs):  # column k of n entries
            self._gilu.apply(gx[k * n:(k + 1) * n], gr[k * n:(k + 1) * n])
        r[...] = gr.cpu().numpy().reshape(r.shape)
        return r


def _to_dev(engine, a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to(f"cuda:{engine.device()}")


def _mixable(fn, outs=()):
    import functools
    import inspect

    sig = inspect.signature(fn)

    @functools.wraps(fn)
    def run(engine, *args, **kw):
        if not isinstance(engine, mixed_engine):
            return fn(engine, *args, **kw)
        ba = sig.bind(engine, *args, **kw)
        moved = {}
        for name, val in list(ba.arguments.items()):
            if isinstance(val, np.ndarray):
                t = _to_dev(engine, val)
                ba.arguments[name] = t
                moved[name] = (val, t)
            elif isinstance(val, mixed_ilu):
                ba.arguments[name] = val.cilu()
        ba.arguments["engine"] = engine.gpu()
        out = fn(*ba.args, **ba.kwargs)
        for name in outs:
            if name in moved:
                host, t = moved[name]
                host[...] = t.cpu().numpy().reshape(host.shape)
        return out

    return run


_make_ilu_gpu = make_ilu


def make_ilu(engine, pntr, indx, vals, policy="N", want_f32=False):  # noqa: F811
    """make_ilu(engine, pntr, indx, vals, policy): gpu/hala_gpu_ilu.hpp:191-195 (gpu_engine) and
    wax/hala_lib_extensions.hpp:368-376 (mixed_engine: host arrays, returns a mixed_ilu)."""
    if isinstance(engine, mixed_engine):
        g = engine.gpu()
        return mixed_ilu(engine, _make_ilu_gpu(g, _to_dev(engine, pntr), _to_dev(engine, indx),
                                               _to_dev(engine, vals), policy, want_f32))
    return _make_ilu_gpu(engine, pntr, indx, vals, policy, want_f32)


dot = _mixable(dot)
norm2 = _mixable(norm2)
axpy = _mixable(axpy, ("y",))
scal = _mixable(scal, ("x",))
vcopy = _mixable(vcopy, ("y",))
sparse_gemv = _mixable(sparse_gemv, ("y",))
make_sparse_matrix = _mixable(make_sparse_matrix)
make_triangular_matrix = _mixable(make_triangular_matrix)
solve_cg = _mixable(solve_cg, ("x", "hist"))
solve_cg_ilu = _mixable(solve_cg_ilu, ("x", "hist"))
solve_batch_cg = _mixable(solve_batch_cg, ("X", "hist"))
solve_batch_cg_ilu = _mixable(solve_batch_cg_ilu, ("X", "hist"))
solve_gmres = _mixable(solve_gmres, ("x", "hist"))
solve_gmres_ilu = _mixable(solve_gmres_ilu, ("x", "hist"))
__all__ += ["mixed_engine", "mixed_ilu"]
