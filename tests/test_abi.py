"""The C ABI library loads and exports every entry point include/hala_amd.h
declares; the product path refuses to run without a GPU (no CPU fallback)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hala_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hala_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) > 60
    for must in ("hala_spmv_f64", "hala_ilu_factor_f64", "hala_cg_f64", "hala_gmres_f64", "hala_gmres_ir_f64",
                 "hala_batch_pcg_f64", "hala_grad_vals_f64", "hala_trsv_f64", "hala_dot_f64"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    from hala_amd import _lib

    L = _lib.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (hala_[a-z0-9_]+)", out))
    assert set(declared_symbols()) <= exported


def test_python_signatures_cover_header():
    from hala_amd import _lib

    missing = [s for s in declared_symbols() if s not in _lib.SIGNATURES]
    assert not missing, missing


def test_no_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from hala_amd import _lib

    h = C.c_void_p()
    rc = _lib.lib().hala_ctx_create(0, None, C.byref(h))
    assert rc != 0
    assert _lib.lib().hala_last_error()


def test_status_codes_and_host_errors():
    from hala_amd import _lib

    L = _lib.lib()
    assert L.hala_ctx_create(0, None, None) == -1  # HALA_ERR_ARG: null output
    n, z = C.c_int(), C.c_int64()
    assert L.hala_gen_laplace2d(0, 5, 4.0, -1.0, C.byref(n), C.byref(z), None, None, None) == -1
    assert L.hala_gen_laplace2d(3, 4, 4.0, -1.0, C.byref(n), C.byref(z), None, None, None) == 0
    assert (n.value, z.value) == (12, 5 * 12 - 2 * 3 - 2 * 4)
