"""CPU oracle (TEST INFRASTRUCTURE ONLY).

ctypes wrapper of oracle/liboracle.so, the C restatement of HALA's CPU
sparse-solve path, and of oracle/_ref/libhala_ref.so, HALA's own BLAS-free CPU
kernels compiled from the reference tree (present only where that tree was).
Only tests, the smoke check and the benchmark's CPU-baseline leg import this.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "libhala_ref.so")

_o = None
_r = None

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f64 = C.c_double
f32 = C.c_float
ch = C.c_char
pi32 = C.POINTER(C.c_int)

_SIG = {
    "hala_or_set_threads": (None, [i32]),
    "hala_or_get_threads": (i32, []),
    "hala_or_cdot": (f64, [i64, vp, vp]),
    "hala_or_cnrm2": (f64, [i64, vp]),
    "hala_or_team_dot": (f64, [i64, vp, vp, i32]),
    "hala_or_spmv": (None, [ch, i32, i32, f64, vp, vp, vp, vp, f64, vp]),
    "hala_or_spmv_f32": (None, [i32, f32, vp, vp, vp, vp, f32, vp]),
    "hala_or_diag": (i32, [i32, vp, vp, vp]),
    "hala_or_levels": (i32, [i32, vp, vp, vp, i32, vp, vp, vp]),
    "hala_or_ilu0": (None, [i32, vp, vp, vp, vp]),
    "hala_or_ilu0_kij": (None, [i32, vp, vp, vp, vp]),
    "hala_or_ilu_apply": (None, [i32, vp, vp, vp, vp, vp]),
    "hala_or_ilu_apply_f32": (None, [i32, vp, vp, vp, vp, vp]),
    "hala_or_ilu_apply_f32_jacobi": (None, [i32, vp, vp, vp, vp, i32, vp, vp]),
    "hala_or_ilu_apply_jacobi": (None, [i32, vp, vp, vp, vp, i32, vp, vp]),
    "hala_or_ilu_apply_t": (None, [i32, vp, vp, vp, vp, vp]),
    "hala_or_trsv": (None, [ch, ch, ch, i32, f64, vp, vp, vp, vp, vp]),
    "hala_or_trsv_f32": (None, [ch, ch, ch, i32, f32, vp, vp, vp, vp, vp]),
    "hala_or_ztrsv_z": (None, [ch, ch, ch, i32, f64, f64, vp, vp, vp, vp, vp]),
    "hala_or_ztrsv_c": (None, [ch, ch, ch, i32, f32, f32, vp, vp, vp, vp, vp]),
    "hala_or_cg": (i32, [i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, f64, i32, vp, i32]),
    "hala_or_cg_team": (i32, [i32, vp, vp, vp, vp, vp, vp, f64, i32, i32, vp, i32]),
    "hala_or_cg_batch": (i32, [i32, vp, vp, vp, i32, vp, vp, vp, vp, i32, vp, vp, f64, i32, vp, i32]),
    "hala_or_gmres": (i32, [i32, vp, vp, vp, i32, vp, vp, vp, vp, f64, i32, i32, vp, i32, pi32]),
    "hala_or_gmres_ir": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, f64, f64, i32, i32, vp, i32, pi32, i32]),
    "hala_or_rotg": (None, [vp, vp, vp, vp]),
    "hala_or_rot": (None, [vp, vp, f64, f64]),
    "hala_or_tpsv_un": (None, [i32, vp, vp]),
    "hala_or_grad_vals": (None, [i32, vp, vp, vp, vp, vp]),
    "hala_or_grad_coef": (None, [i32, vp, vp, vp, vp, vp]),
    "hala_or_transpose": (None, [i32, i32, vp, vp, vp, vp, vp, vp]),
    "hala_or_coo_to_csr": (i64, [i32, i32, i64, vp, vp, vp, i32, vp, vp, vp]),
    "hala_or_zdot_z": (None, [i64, vp, vp, i32, vp]),
    "hala_or_zdot_c": (None, [i64, vp, vp, i32, vp]),
    "hala_or_znrm2_z": (f64, [i64, vp]),
    "hala_or_znrm2_c": (f32, [i64, vp]),
    "hala_or_zspmv_z": (None, [ch, i32, i32, f64, f64, vp, vp, vp, vp, f64, f64, vp]),
    "hala_or_zspmv_c": (None, [ch, i32, i32, f32, f32, vp, vp, vp, vp, f32, f32, vp]),
    "hala_or_zilu0_z": (None, [i32, vp, vp, vp, vp]),
    "hala_or_zilu0_c": (None, [i32, vp, vp, vp, vp]),
    "hala_or_zilu_apply_z": (None, [i32, vp, vp, vp, vp, vp, vp]),
    "hala_or_zilu_apply_c": (None, [i32, vp, vp, vp, vp, vp, vp]),
}
_REF_SIG = {
    "hala_ref_diag": [i32, vp, vp, vp],
    "hala_ref_ilu0": [i32, vp, vp, vp, vp],
    "hala_ref_ilu_apply": [i32, vp, vp, vp, vp, vp],
    "hala_ref_spmv_n": [i32, i32, f64, vp, vp, vp, vp, f64, vp],
    "hala_ref_spmv_n_f32": [i32, i32, f32, vp, vp, vp, vp, f32, vp],
    "hala_ref_ilu_apply_f32": [i32, vp, vp, vp, vp, vp],
}

PREC_IDENTITY, PREC_ILU, PREC_JACOBI, PREC_ILU_T = 0, 1, 2, 3


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        for k, (res, args) in _SIG.items():
            fn = getattr(L, k)
            fn.restype = res
            fn.argtypes = args
        _o = L
    return _o


def ref():
    """HALA's own kernels, or None where the reference tree was not available to build them."""
    global _r
    if _r is None and os.path.exists(REF):
        L = C.CDLL(REF)
        for k, args in _REF_SIG.items():
            fn = getattr(L, k)
            fn.restype = None
            fn.argtypes = args
        _r = L
    return _r


def P(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def ci32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---- thin numpy-friendly wrappers ---------------------------------------------
def set_threads(t):
    """Threads of the oracle (1 by default); every result is bit-identical for any count."""
    lib().hala_or_set_threads(int(t))


def get_threads():
    return lib().hala_or_get_threads()


def host_threads(cap=16):
    """The CPU threads this process may use, at most cap."""
    try:
        k = len(os.sched_getaffinity(0))
    except AttributeError:
        k = os.cpu_count() or 1
    return max(1, min(cap, k))


def cdot(x, y):
    x, y = c64(x), c64(y)
    return lib().hala_or_cdot(x.size, P(x), P(y))


def cnrm2(x):
    x = c64(x)
    return lib().hala_or_cnrm2(x.size, P(x))


def team_dot(x, y, team):
    x, y = c64(x), c64(y)
    return lib().hala_or_team_dot(x.size, P(x), P(y), team)


def spmv(trans, nrows, ncols, alpha, pntr, indx, vals, x, beta, y):
    y = c64(y).copy()
    lib().hala_or_spmv(trans.encode(), nrows, ncols, alpha, P(ci32(pntr)), P(ci32(indx)), P(c64(vals)), P(c64(x)),
                       beta, P(y))
    return y


def spmv_f32(pntr, indx, vals32, x32, alpha=1.0, beta=0.0, y=None):
    n = len(pntr) - 1
    y = np.zeros(n, np.float32) if y is None else np.ascontiguousarray(y, np.float32).copy()
    lib().hala_or_spmv_f32(n, alpha, P(ci32(pntr)), P(ci32(indx)), P(np.ascontiguousarray(vals32, np.float32)),
                           P(np.ascontiguousarray(x32, np.float32)), beta, P(y))
    return y


# ---- complex scalars (numpy complex128 / complex64 arrays) -------------------
def _zs(a):
    """(array, suffix): complex128 -> 'z', complex64 -> 'c'."""
    a = np.ascontiguousarray(a)
    if a.dtype == np.complex128:
        return a, "z"
    if a.dtype == np.complex64:
        return a, "c"
    raise TypeError("complex128 or complex64 expected")


def zdot(x, y, conj=True):
    """sum conj(x_i) y_i (conj) or sum x_i y_i, canonical order per part."""
    x, sf = _zs(x)
    y = np.ascontiguousarray(y, dtype=x.dtype)
    res = np.zeros(2, np.float64 if sf == "z" else np.float32)
    getattr(lib(), "hala_or_zdot_" + sf)(x.size, P(x), P(y), int(bool(conj)), P(res))
    return complex(res[0], res[1])


def znrm2(x):
    x, sf = _zs(x)
    return getattr(lib(), "hala_or_znrm2_" + sf)(x.size, P(x))


def zspmv(trans, nrows, ncols, alpha, pntr, indx, vals, x, beta, y):
    vals, sf = _zs(vals)
    x = np.ascontiguousarray(x, dtype=vals.dtype)
    getattr(lib(), "hala_or_zspmv_" + sf)(trans.encode() if isinstance(trans, str) else trans, nrows, ncols,
                                           alpha.real, alpha.imag, P(ci32(pntr)), P(ci32(indx)), P(vals), P(x),
                                           beta.real, beta.imag, P(y))
    return y


def zilu0(pntr, indx, vals, dg):
    vals, sf = _zs(vals)
    f = vals.copy()
    getattr(lib(), "hala_or_zilu0_" + sf)(pntr.size - 1, P(ci32(pntr)), P(ci32(indx)), P(ci32(dg)), P(f))
    return f


def zilu_apply(pntr, indx, dg, f, x):
    f, sf = _zs(f)
    x = np.ascontiguousarray(x, dtype=f.dtype)
    r = np.zeros_like(x)
    getattr(lib(), "hala_or_zilu_apply_" + sf)(pntr.size - 1, P(ci32(pntr)), P(ci32(indx)), P(ci32(dg)), P(f), P(x),
                                                P(r))
    return r


def coo_to_csr(nrows, ncols, rows, cols, vals, add_diag=False):
    """Host CSR (pntr, indx, vals) of COO triplets: (row, column, position) order, duplicates summed
    left to right, optional zero diagonal inserted.  Raises ValueError for an entry out of range."""
    rows, cols, vals = ci32(rows), ci32(cols), c64(vals)
    pntr = np.zeros(nrows + 1, np.int32)
    nnz = lib().hala_or_coo_to_csr(nrows, ncols, rows.size, P(rows), P(cols), P(vals), int(bool(add_diag)), P(pntr),
                                   None, None)
    if nnz < 0:
        raise ValueError(f"entry {-1 - nnz} out of range")
    indx = np.zeros(nnz, np.int32)
    out = np.zeros(nnz, np.float64)
    lib().hala_or_coo_to_csr(nrows, ncols, rows.size, P(rows), P(cols), P(vals), int(bool(add_diag)), P(pntr),
                             P(indx), P(out))
    return pntr, indx, out


def diag(pntr, indx):
    n = len(pntr) - 1
    d = np.empty(n, np.int32)
    bad = lib().hala_or_diag(n, P(ci32(pntr)), P(ci32(indx)), P(d))
    return d, bad


def levels(pntr, indx, dg, upper):
    n = len(pntr) - 1
    lev = np.empty(n, np.int32)
    lptr = np.empty(n + 1, np.int32)
    perm = np.empty(n, np.int32)
    nl = lib().hala_or_levels(n, P(ci32(pntr)), P(ci32(indx)), P(ci32(dg)), int(upper), P(lev), P(lptr), P(perm))
    return lev, lptr[: nl + 1].copy(), perm


def ilu0(pntr, indx, vals, dg=None, kij=False):
    n = len(pntr) - 1
    if dg is None:
        dg, _ = diag(pntr, indx)
    f = c64(vals).copy()
    fn = lib().hala_or_ilu0_kij if kij else lib().hala_or_ilu0
    fn(n, P(ci32(pntr)), P(ci32(indx)), P(ci32(dg)), P(f))
    return f


def ilu_apply(pntr, indx, dg, ilu, x):
    x = c64(x).copy()
    lib().hala_or_ilu_apply(len(pntr) - 1, P(ci32(pntr)), P(ci32(indx)), P(ci32(dg)), P(c64(ilu)), P(x))
    return x


def ilu_apply_f32(pntr, indx, dg, ilu32, x32):
    x = np.ascontiguousarray(x32, np.float32).copy()
    lib().hala_or_ilu_apply_f32(len(pntr) - 1, P(ci32(pntr)), P(ci32(indx)), P(ci32(dg)),
                                P(np.ascontiguousarray(ilu32, np.float32)), P(x))
    return x


def ilu_apply_f32_jacobi(pntr, indx, dg, ilu32, x32, k):
    """(LU)^-1 x with the FP32 factor applied by k Jacobi sweeps per triangle (hala_or_ilu_apply_f32_jacobi)."""
    x = np.ascontiguousarray(x32, np.float32).copy()
    w = np.empty(3 * x.size, np.float32)
    lib().hala_or_ilu_apply_f32_jacobi(len(pntr) - 1, P(ci32(pntr)), P(ci32(indx)), P(ci32(dg)),
                                       P(np.ascontiguousarray(ilu32, np.float32)), int(k), P(x), P(w))
    return x


def ilu_apply_jacobi(pntr, indx, dg, ilu, x, k):
    """(LU)^-1 x with the FP64 factor applied by k Jacobi sweeps per triangle (hala_or_ilu_apply_jacobi)."""
    x = c64(x).copy()
    w = np.empty(3 * x.size, np.float64)
    lib().hala_or_ilu_apply_jacobi(len(pntr) - 1, P(ci32(pntr)), P(ci32(indx)), P(ci32(dg)), P(c64(ilu)), int(k),
                                   P(x), P(w))
    return x


def ilu_apply_t(pntr, indx, dg, ilu, x):
    x = c64(x).copy()
    lib().hala_or_ilu_apply_t(len(pntr) - 1, P(ci32(pntr)), P(ci32(indx)), P(ci32(dg)), P(c64(ilu)), P(x))
    return x


def trsv(uplo, trans, dg, alpha, pntr, indx, vals, b):
    n = len(pntr) - 1
    x = np.zeros(n, np.float64)
    lib().hala_or_trsv(uplo.encode(), trans.encode(), dg.encode(), n, alpha, P(ci32(pntr)), P(ci32(indx)),
                       P(c64(vals)), P(c64(b)), P(x))
    return x


def trsv_f32(uplo, trans, dg, alpha, pntr, indx, vals, b):
    """FP32 values, vectors and arithmetic, the term order of trsv()."""
    n = len(pntr) - 1
    x = np.zeros(n, np.float32)
    lib().hala_or_trsv_f32(uplo.encode(), trans.encode(), dg.encode(), n, float(alpha), P(ci32(pntr)), P(ci32(indx)),
                           P(np.ascontiguousarray(vals, np.float32)), P(np.ascontiguousarray(b, np.float32)), P(x))
    return x


def ztrsv(uplo, trans, dg, alpha, pntr, indx, vals, b):
    """Complex triangular solve; complex128 inputs use the double version, complex64 the float one."""
    vals = np.ascontiguousarray(vals)
    sf = "z" if vals.dtype == np.complex128 else "c"
    b = np.ascontiguousarray(b, vals.dtype)
    n = len(pntr) - 1
    x = np.zeros(n, vals.dtype)
    alpha = complex(alpha)
    getattr(lib(), "hala_or_ztrsv_" + sf)(uplo.encode(), trans.encode(), dg.encode(), n, alpha.real, alpha.imag,
                                          P(ci32(pntr)), P(ci32(indx)), P(vals), P(b), P(x))
    return x


def cg(pntr, indx, vals, b, x0=None, tol=0.0, max_iter=2**31 - 1, prec=PREC_IDENTITY, mpntr=None, mindx=None,
       mdiag=None, mdata=None, max_hist=100000):
    n = len(pntr) - 1
    x = np.zeros(n) if x0 is None else c64(x0).copy()
    hist = np.zeros(max_hist)
    it = lib().hala_or_cg(n, P(ci32(pntr)), P(ci32(indx)), P(c64(vals)), prec,
                          P(ci32(mpntr)) if mpntr is not None else None,
                          P(ci32(mindx)) if mindx is not None else None,
                          P(ci32(mdiag)) if mdiag is not None else None,
                          P(c64(mdata)) if mdata is not None else None, P(c64(b)), P(x), tol, max_iter, P(hist),
                          max_hist)
    return it, x, hist[: max(it - 1, 0)].copy()


def cg_batch(pntr, indx, vals, B, tol=0.0, max_iter=2**31 - 1, prec=PREC_IDENTITY, mdiag=None, mdata=None,
             max_hist=100000):
    """Multi-RHS PCG with a shared matrix (hala_or_cg_batch); B is (nrhs, n): row k is right-hand side k."""
    B = np.ascontiguousarray(B, np.float64)
    nrhs, n = B.shape
    X = np.zeros_like(B)
    hist = np.zeros(max_hist)
    ilu_ = prec in (PREC_ILU,)
    it = lib().hala_or_cg_batch(n, P(ci32(pntr)), P(ci32(indx)), P(c64(vals)), prec,
                                P(ci32(pntr)) if ilu_ else None, P(ci32(indx)) if ilu_ else None,
                                P(ci32(mdiag)) if mdiag is not None else None,
                                P(c64(mdata)) if mdata is not None else None, nrhs, P(B), P(X), tol, max_iter,
                                P(hist), max_hist)
    return it, X, hist[: max(it - 1, 0)].copy()


def cg_team(pntr, indx, vals, dvals, b, tol, team, max_iter=2**31 - 1, max_hist=100000):
    n = len(pntr) - 1
    x = np.zeros(n)
    hist = np.zeros(max_hist)
    it = lib().hala_or_cg_team(n, P(ci32(pntr)), P(ci32(indx)), P(c64(vals)), P(c64(dvals)), P(c64(b)), P(x), tol,
                               max_iter, team, P(hist), max_hist)
    return it, x, hist[: max(it - 1, 0)].copy()


def gmres(pntr, indx, vals, b, x0=None, tol=0.0, restart=30, max_outer=2**31 - 1, prec=PREC_ILU, dg=None,
          ilu=None, max_hist=100000):
    n = len(pntr) - 1
    x = np.zeros(n) if x0 is None else c64(x0).copy()
    hist = np.zeros(max_hist)
    nh = C.c_int(0)
    tot = lib().hala_or_gmres(n, P(ci32(pntr)), P(ci32(indx)), P(c64(vals)), prec,
                              P(ci32(dg)) if dg is not None else None, P(c64(ilu)) if ilu is not None else None,
                              P(c64(b)), P(x), tol, restart, max_outer, P(hist), max_hist, C.byref(nh))
    return tot, x, hist[: nh.value].copy()


def gmres_ir(pntr, indx, vals, dg, ilu64, b, rtol=1e-10, inner_rtol=1e-4, restart=30, max_refine=50, fp32=True,
             jacobi=0):
    """Iterative refinement around ILU(0)-GMRES; fp32 selects the FP32 product and ILU inside;
    jacobi > 0 applies the inner ILU (FP32 or FP64) by Jacobi sweeps."""
    n = len(pntr) - 1
    x = np.zeros(n)
    hist = np.zeros(max_refine + 2)
    nh = C.c_int(0)
    v32 = np.ascontiguousarray(vals, np.float32) if fp32 else None
    i32_ = np.ascontiguousarray(ilu64, np.float32) if fp32 else None
    tot = lib().hala_or_gmres_ir(n, P(ci32(pntr)), P(ci32(indx)), P(c64(vals)), P(v32), P(ci32(dg)), P(c64(ilu64)),
                                 P(i32_), P(c64(b)), P(x), rtol, inner_rtol, restart, max_refine, P(hist), hist.size,
                                 C.byref(nh), int(jacobi))
    return tot, x, hist[: nh.value].copy()


def gmres_ir32(pntr, indx, vals, dg, ilu64, b, **kw):
    return gmres_ir(pntr, indx, vals, dg, ilu64, b, fp32=True, **kw)


def rotg(a, b):
    a_, b_, c_, s_ = (C.c_double(a), C.c_double(b), C.c_double(), C.c_double())
    lib().hala_or_rotg(C.byref(a_), C.byref(b_), C.byref(c_), C.byref(s_))
    return a_.value, b_.value, c_.value, s_.value


def tpsv_un(ap, x):
    x = c64(x).copy()
    lib().hala_or_tpsv_un(x.size, P(c64(ap)), P(x))
    return x


def grad_vals(pntr, indx, lam, x):
    g = np.empty(len(indx), np.float64)
    lib().hala_or_grad_vals(len(pntr) - 1, P(ci32(pntr)), P(ci32(indx)), P(c64(lam)), P(c64(x)), P(g))
    return g


def grad_coef(cptr, cnnz, ccoef, gvals):
    nc = len(cptr) - 1
    g = np.empty(nc, np.float64)
    lib().hala_or_grad_coef(nc, P(ci32(cptr)), P(ci32(cnnz)), P(c64(ccoef)), P(c64(gvals)), P(g))
    return g


def transpose(nrows, ncols, pntr, indx, vals):
    nnz = len(indx)
    tp = np.empty(ncols + 1, np.int32)
    ti = np.empty(nnz, np.int32)
    tv = np.empty(nnz, np.float64)
    lib().hala_or_transpose(nrows, ncols, P(ci32(pntr)), P(ci32(indx)), P(c64(vals)), P(tp), P(ti), P(tv))
    return tp, ti, tv
