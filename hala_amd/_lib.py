"""ctypes binding of libhala_amd.so (the C ABI declared in include/hala_amd.h).

The product path: every call lands in the gfx950 library.  There is no CPU
fallback -- if the shared library is missing the import of a compute entry
point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhala_amd.so")

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
u64 = C.c_uint64
f64 = C.c_double
f32 = C.c_float
ch = C.c_char
pi32 = C.POINTER(C.c_int)
pi64 = C.POINTER(C.c_int64)
pf64 = C.POINTER(C.c_double)
pf32 = C.POINTER(C.c_float)
pvp = C.POINTER(C.c_void_p)


class CsrInfo(C.Structure):
    _fields_ = [("nrows", i32), ("ncols", i32), ("nnz", i64), ("max_row", i32), ("min_row", i32),
                ("mean_row", f64), ("kernel", i32), ("index_bytes", i32)]


class IluInfo(C.Structure):
    _fields_ = [("n", i32), ("nnz", i64), ("nlev_lower", i32), ("nlev_upper", i32), ("nchunk_lower", i32),
                ("nchunk_upper", i32), ("slots_lower", i64), ("slots_upper", i64)]


# name -> argtypes (restype is int status unless listed in _RESTYPE)
SIGNATURES = {
    "hala_last_error": [],
    "hala_version": [pi32, pi32],
    "hala_device_count": [pi32],
    "hala_ctx_create": [i32, vp, pvp],
    "hala_ctx_destroy": [vp],
    "hala_ctx_set_stream": [vp, vp],
    "hala_ctx_sync": [vp],
    "hala_ctx_device": [vp, pi32],
    "hala_ctx_set_option": [vp, C.c_char_p, i32],
    "hala_ctx_get_option": [vp, C.c_char_p, pi32],
    "hala_malloc": [vp, C.c_size_t, pvp],
    "hala_free": [vp, vp],
    "hala_memcpy": [vp, vp, vp, C.c_size_t, i32],
    "hala_fill_f64": [vp, vp, i64, f64],
    "hala_fill_f32": [vp, vp, i64, f32],
    "hala_fill_i32": [vp, vp, i64, i32],
    "hala_convert_f64_f32": [vp, i64, vp, vp],
    "hala_convert_f32_f64": [vp, i64, vp, vp],
    "hala_dot_f64": [vp, i64, vp, vp, pf64],
    "hala_nrm2_f64": [vp, i64, vp, pf64],
    "hala_axpy_f64": [vp, i64, f64, vp, vp],
    "hala_scal_f64": [vp, i64, f64, vp],
    "hala_copy_f64": [vp, i64, vp, vp],
    "hala_dot_f32": [vp, i64, vp, vp, pf32],
    "hala_nrm2_f32": [vp, i64, vp, pf32],
    "hala_axpy_f32": [vp, i64, f32, vp, vp],
    "hala_scal_f32": [vp, i64, f32, vp],
    "hala_copy_f32": [vp, i64, vp, vp],
    "hala_vdivide_f64": [vp, i64, vp, vp, vp],
    "hala_gemv_f64": [vp, ch, i32, i32, f64, vp, i32, vp, f64, vp],
    "hala_csr_create": [vp, i32, i32, i64, vp, vp, pvp],
    "hala_csr_destroy": [vp],
    "hala_csr_info": [vp, C.POINTER(CsrInfo)],
    "hala_csr_set_kernel": [vp, i32],
    "hala_csr_validate": [vp, i32, pi32],
    "hala_spmv_f64": [vp, ch, f64, vp, vp, f64, vp],
    "hala_spmv_f32": [vp, ch, f32, vp, vp, f32, vp],
    "hala_spmv_f32x64": [vp, f64, vp, vp, f64, vp],
    "hala_ilu_create": [vp, i32, i64, vp, vp, pvp],
    "hala_ilu_destroy": [vp],
    "hala_ilu_info": [vp, C.POINTER(IluInfo)],
    "hala_ilu_sweep_info": [vp, i32, pi32],
    "hala_ilu_stencil_info": [vp, pi32],
    "hala_ilu_factor_f64": [vp, vp, i32],
    "hala_ilu_apply_f64": [vp, vp, vp],
    "hala_ilu_apply_f32": [vp, vp, vp],
    "hala_ilu_apply_f32x64": [vp, vp, vp],
    "hala_ilu_apply_t_f64": [vp, vp, vp],
    "hala_ilu_get_symbolic": [vp, i32, vp, vp, vp, vp],
    "hala_ilu_get_factor_f64": [vp, vp],
    "hala_tri_create": [vp, ch, ch, i32, i64, vp, vp, pvp],
    "hala_tri_destroy": [vp],
    "hala_trsv_f64": [vp, ch, f64, vp, vp, vp],
    "hala_trsv_f32": [vp, ch, f32, vp, vp, vp],
    "hala_spmm_f64": [vp, ch, ch, i32, i32, f64, vp, vp, i32, f64, vp, i32],
    "hala_spmm_f32": [vp, ch, ch, i32, i32, f32, vp, vp, i32, f32, vp, i32],
    "hala_spmm_z": [vp, ch, ch, i32, i32, f64, f64, vp, vp, i32, f64, f64, vp, i32],
    "hala_spmm_c": [vp, ch, ch, i32, i32, f32, f32, vp, vp, i32, f32, f32, vp, i32],
    "hala_trsv_z": [vp, ch, f64, f64, vp, vp, vp],
    "hala_trsv_c": [vp, ch, f32, f32, vp, vp, vp],
    "hala_trsm_f64": [vp, ch, ch, i32, f64, vp, vp, i32],
    "hala_trsm_f32": [vp, ch, ch, i32, f32, vp, vp, i32],
    "hala_trsm_z": [vp, ch, ch, i32, f64, f64, vp, vp, i32],
    "hala_trsm_c": [vp, ch, ch, i32, f32, f32, vp, vp, i32],
    "hala_cg_f64": [vp, vp, i32, vp, vp, vp, vp, f64, i32, pi32, vp, i32],
    "hala_cg_f64_noalloc": [vp, vp, i32, vp, vp, vp, vp, f64, i32, pi32, vp, i32],
    "hala_cg_t_f64": [vp, vp, vp, vp, vp, f64, i32, pi32, vp, i32],
    "hala_cg_batch_f64": [vp, vp, i32, vp, i32, vp, vp, f64, i32, pi32, vp, i32],
    "hala_gmres_f64": [vp, vp, i32, vp, vp, vp, f64, i32, i32, pi32, vp, i32, pi32],
    "hala_gmres_ir_f64": [vp, vp, vp, vp, vp, vp, f64, f64, i32, i32, pi32, vp, i32, pi32],
    "hala_assemble_f64": [vp, i64, vp, vp, vp, vp, vp],
    "hala_batch_create": [vp, i32, vp, vp, vp, pvp],
    "hala_batch_destroy": [vp],
    "hala_batch_set_shape": [vp, i32, i32],
    "hala_batch_get_shape": [vp, pi32, pi32],
    "hala_batch_pcg_f64": [vp, vp, vp, vp, vp, f64, i32, vp],
    "hala_batch_tune": [vp, vp, vp, vp, vp, f64, i32, i32, pi32, pi32, pf64],
    "hala_grad_vals_f64": [vp, vp, vp, vp],
    "hala_grad_coef_f64": [vp, i32, vp, vp, vp, vp, vp],
    "hala_gen_laplace2d": [i32, i32, f64, f64, pi32, pi64, vp, vp, vp],
    "hala_gen_laplace3d": [i32, i32, i32, pi32, pi64, vp, vp, vp],
    "hala_gen_discretize_pde": [i32, pi32, pi64, vp, vp, vp, vp],
    "hala_gen_convdiff27": [i32, u64, f64, pi32, pi64, vp, vp, vp],
    "hala_gen_uniform": [u64, i64, f64, f64, vp],
    "hala_gen_batch_laplace": [i32, u64, i32, i32, pi32, pi64, vp, vp, vp, vp, vp, vp, vp],
    "hala_gen_diffusion3d": [i32, u64, pi32, pi64, pi32, pi64, vp, vp, vp, vp, vp, vp, vp],
    "hala_gen_gaussian_source": [i32, f64, vp],
    "hala_mm_info": [C.c_char_p, pi32, pi32, pi64, pi64, pi32, pi32],
    "hala_mm_read_coo": [C.c_char_p, i32, i64, vp, vp, vp, pi64],
    "hala_mm_write_csr": [C.c_char_p, i32, i32, vp, vp, vp],
    "hala_coo_to_csr": [vp, i32, i32, i64, vp, vp, vp, i32, vp, vp, vp, pi64],
    "hala_dotc_z": [vp, i64, vp, vp, vp],
    "hala_dotu_z": [vp, i64, vp, vp, vp],
    "hala_nrm2_z": [vp, i64, vp, vp],
    "hala_axpy_z": [vp, i64, f64, f64, vp, vp],
    "hala_scal_z": [vp, i64, f64, f64, vp],
    "hala_spmv_z": [vp, ch, f64, f64, vp, vp, f64, f64, vp],
    "hala_ilu_factor_z": [vp, vp],
    "hala_ilu_apply_z": [vp, vp, vp],
    "hala_dotc_c": [vp, i64, vp, vp, vp],
    "hala_dotu_c": [vp, i64, vp, vp, vp],
    "hala_nrm2_c": [vp, i64, vp, vp],
    "hala_axpy_c": [vp, i64, f32, f32, vp, vp],
    "hala_scal_c": [vp, i64, f32, f32, vp],
    "hala_spmv_c": [vp, ch, f32, f32, vp, vp, f32, f32, vp],
    "hala_ilu_factor_c": [vp, vp],
    "hala_ilu_apply_c": [vp, vp, vp],
}
_RESTYPE = {"hala_last_error": C.c_char_p}

_lib = None


class HalaError(RuntimeError):
    """A non-zero status from libhala_amd (HALA raises std::runtime_error for library errors)."""

    def __init__(self, code, msg):
        super().__init__(f"libhala_amd error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run python hala_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def call(name, *args):
    """Calls an entry point and raises HalaError on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise HalaError(rc, lib().hala_last_error().decode(errors="replace"))
    return rc
