"""ctypes binding of libbkx.so, the gfx950 kernels behind include/bkx.h.

There is no fallback: when the library is missing, or a call reports an
error, this module raises.  Build the library with ``make -C omp_bk_amd``
(``__graft_entry__.build()`` does the same)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbkx.so")

_c_int = ctypes.c_int
_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p

# name -> argument types (all return int status)
_SIGNATURES = {
    "bkx_bk1_apply_f64": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk3_apply_f64": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk5_apply_f64": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk5_apply_f32c": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk5_apply_f32w": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk5_apply_bf16s": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk5_apply_oz": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk1_apply_ff": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk1_apply_f32c": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk1_apply_f32w": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk3_apply_ff": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk3_apply_f32c": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk3_apply_f32w": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk3_basis_prepare": [_c_int, _vp, _vp, _vp],
    "bkx_bk3_apply_prepared_f64": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk3_apply_prepared_f32w": [_c_int, _c_size, _vp, _vp, _vp, _vp, _vp],
    "bkx_bk3_adjoint_prepared_f64": [_c_int, _c_size] + [_vp] * 9,
    "bkx_bk3_adjoint_prepared_f32w": [_c_int, _c_size] + [_vp] * 9,
    "bkx_bk3_apply_kappa_f64": [_c_int, _c_size] + [_vp] * 7,
    "bkx_bk3_adjoint_f64": [_c_int, _c_size] + [_vp] * 10,
    "bkx_bk3_vjp_f64": [_c_int, _c_size] + [_vp] * 9,
    "bkx_bk3_apply_kappa_f32w": [_c_int, _c_size] + [_vp] * 7,
    "bkx_bk3_adjoint_f32w": [_c_int, _c_size] + [_vp] * 10,
    "bkx_bk3_vjp_f32w": [_c_int, _c_size] + [_vp] * 9,
    "bkx_bk3_apply_kappa_f32c": [_c_int, _c_size] + [_vp] * 7,
    "bkx_bk3_adjoint_f32c": [_c_int, _c_size] + [_vp] * 10,
    "bkx_bk3_vjp_f32c": [_c_int, _c_size] + [_vp] * 9,
    "bkx_sum_f64": [_vp, _c_size, _vp, _vp],
    "bkx_bk3_adjoint_loss_f64": [_c_int, _c_size] + [_vp] * 10,
    "bkx_bk5_make_G_f64": [_c_int, _c_size, _vp, _c_int, _c_int, _vp, _vp],
    "bkx_bk5_make_G_f32": [_c_int, _c_size, _vp, _c_int, _c_int, _vp, _vp],
    "bkx_bk5_interleave_G_f64": [_c_int, _c_size, _c_int, _c_int, _vp, _vp, _vp],
    "bkx_bk5_interleave_G_f32": [_c_int, _c_size, _c_int, _c_int, _vp, _vp, _vp],
    "bkx_bk1_transpose_jxw_f64": [_c_int, _c_size, _vp, _vp, _vp],
    "bkx_bk1_transpose_jxw_f32": [_c_int, _c_size, _vp, _vp, _vp],
    "bkx_geo_bk3_f64": [_c_int, _c_size] + [_vp] * 7,
    "bkx_geo_bk5_f64": [_c_int, _c_size] + [_vp] * 5,
    "bkx_restrict_create": [_c_size, _c_int, _vp, _c_size, _vp],
    "bkx_restrict_destroy": [_vp],
    "bkx_restrict_info": [_vp, _vp, _vp, _vp],
    "bkx_restrict_gather_f64": [_vp, _vp, _vp, _vp],
    "bkx_restrict_scatter_f64": [_vp, _vp, _vp, _vp],
    "bkx_bk5_apply_restricted_f64": [_vp, _c_int, _vp, _vp, _vp, _vp, _vp],
    "bkx_probe_copy": [_vp, _vp, _c_size, _vp],
    "bkx_probe_read_mix": [_c_int, _vp, _vp, _c_size, _vp],
    "bkx_probe_fp64_rate": [_c_int, _c_int, _vp, _c_size, _vp],
    "bkx_probe_read": [_vp, _c_size, _vp, _vp],
    "bkx_tune_count": [_c_int, _c_int],
    "bkx_tune_set": [_c_int, _c_int, _c_int, _c_int],
    "bkx_tune_get": [_c_int, _c_int, _c_int],
    "bkx_min_order": [_c_int],
    "bkx_max_order": [_c_int],
    "bkx_version": [],
}


class BkxError(RuntimeError):
    """A non-zero bkx_status from the C ABI."""

    def __init__(self, fn, status, msg):
        super().__init__("%s failed: %s (status %d)" % (fn, msg, status))
        self.status = status


_lib = None


def lib():
    """The loaded libbkx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libbkx.so not found at %s: build it with `make -C omp_bk_amd`" % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        handle.bkx_status_string.restype = ctypes.c_char_p
        handle.bkx_status_string.argtypes = [_c_int]
        handle.bkx_bk3_basis_size.restype = _c_size
        handle.bkx_bk3_basis_size.argtypes = [_c_int]
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = _c_int
        _lib = handle
    return _lib


def exported_symbols():
    """Names of the C-ABI functions this binding declares."""
    return sorted(list(_SIGNATURES) + ["bkx_status_string", "bkx_bk3_basis_size"])


def call(name, *args):
    """Call a C-ABI entry point; raise BkxError on a non-zero status."""
    h = lib()
    status = getattr(h, name)(*args)
    if status != 0:
        raise BkxError(name, status, h.bkx_status_string(status).decode())
    return status
