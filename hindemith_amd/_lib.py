"""Loader for libhmx.so (the gfx950 backend) with ctypes signatures.

There is no CPU fallback: if the library is missing or does not load, every
entry point raises ``HindemithError``.
"""
from __future__ import annotations

import ctypes
import os

__all__ = ["HindemithError", "lib", "check", "LIB_PATH"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhmx.so")

HMX_OK = 0
_STATUS = {1: "invalid argument", 2: "HIP error", 3: "unsupported", 4: "not found",
           5: "bad state", 6: "timeout"}


class HindemithError(RuntimeError):
    """Error raised for every failure of the backend or the front end."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


c_int, c_int64, c_float, c_double = ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_double
vp = ctypes.c_void_p
c_char_p = ctypes.c_char_p
pint = ctypes.POINTER(ctypes.c_int)
pint64 = ctypes.POINTER(ctypes.c_int64)

# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "hmx_init": [c_int],
    "hmx_compile": [c_int, c_char_p, c_char_p, c_char_p],
    "hmx_bind": [c_int, c_char_p, c_int, vp, c_int64],
    "hmx_run": [c_int],
    "hmx_result_count": [c_int, pint],
    "hmx_result_info": [c_int, c_int, pint, pint, pint, pint64, ctypes.POINTER(c_char_p)],
    "hmx_result_read": [c_int, c_int, vp, c_int64],
    "hmx_plan_text": [c_int, ctypes.c_char_p, c_int64],
    "hmx_program_source": [c_int, c_int, ctypes.c_char_p, c_int64],
    "hmx_release": [c_int],
    "hmx_stencil2d": [vp, vp, c_int, c_int, c_int, vp, vp, vp, vp],
    "hmx_dia2d_spmv": [vp, c_int, vp, vp, vp, vp, c_int, c_int, vp],
    "hmx_csr_spmv": [c_int, c_int, c_int64, vp, vp, vp, vp, vp, c_int, vp],
    "hmx_csr_bin_rows": [c_int, vp, vp, vp],
    "hmx_csr_spmv_binned": [c_int, vp, vp, vp, vp, vp, vp, vp, vp],
    "hmx_csr_row_blocks": [c_int, vp, vp, vp],
    "hmx_csr_spmv_rowblocks": [c_int, vp, vp, vp, vp, vp, vp, c_int, vp],
    "hmx_csr_slice_count": [],
    "hmx_csr_slice_build": [c_int, c_int, vp, vp, vp, vp, vp],
    "hmx_csr_slice_values": [c_int64, vp, vp, vp, vp],
    "hmx_csr_spmv_sliced": [c_int, vp, vp, vp, vp, vp, vp, vp],
    "hmx_sum2d": [vp, c_int64, vp, vp, vp],
    "hmx_hs_setup": [vp, vp, c_int, c_int] + [c_int, vp, vp, vp] * 3 + [vp, vp, vp, vp],
    "hmx_hs_jacobi": [vp, vp, vp, vp, c_float, c_int, c_int, c_int, vp, vp, vp, c_int,
                      vp, vp, vp, vp, vp],
    "hmx_hs_jacobi_cfg": [vp, vp, vp, vp, c_float, c_int, c_int, c_int, vp, vp, vp, c_int,
                          vp, vp, vp, vp, c_int, c_int, vp],
    "hmx_hs_variant_count": [],
    "hmx_hs_variant_info": [c_int, pint, pint, pint, pint],
    "hmx_hs_variant_strip": [c_int],
    "hmx_hs_config": [c_int, c_int, pint, pint],
    "hmx_hs_recip_check": [c_int, c_int, vp, vp],
    "hmx_hs_tune": [vp, vp, vp, c_float, c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp,
                    vp, pint, pint],
    "hmx_hs_solve": [vp, vp, c_int, c_int, c_float] + [c_int, vp, vp, vp] * 4 +
                    [c_int, vp, vp, vp, vp],
    "hmx_hs_batched_variant_count": [],
    "hmx_hs_batched_variant_info": [c_int, pint, pint, pint, pint],
    "hmx_hs_batched_resident": [c_int, c_int, pint, pint, pint],
    "hmx_hs_precondition": [vp, vp, vp, vp, vp, vp, vp, c_int, c_int, c_int, vp, vp],
    "hmx_hs_matrix_powers": [c_float, vp, vp, vp, vp, vp, vp, c_int, c_int, c_int, vp],
    "hmx_warp_img2d": [vp, vp, vp, vp, c_int, c_int, vp],
    "hmx_dia_spmv": [vp, c_int, vp, vp, vp, c_int, vp],
    "hmx_brox_set_matrix": [vp, c_int, vp, c_float, c_int, c_int, vp],
    "hmx_brox_set_redblack_matrix": [vp, vp, c_int, vp, c_float, c_int, c_int, vp],
    "hmx_brox_precondition": [vp, c_int, vp, vp, vp, vp, vp, vp, vp, c_int, c_int, vp],
    "hmx_hs_batched_release": [],
    "hmx_hs_ckpt_count": [c_int, c_int],
    "hmx_hs_forward_ckpt": [vp, vp, c_int, c_int, c_int, vp] + [c_int, vp, vp, vp] * 4 +
                           [c_int, c_int, vp, vp, vp, vp, vp, vp],
    "hmx_hs_adjoint": [vp, vp, c_int, c_int, c_int, vp] + [c_int, vp, vp, vp] * 4 +
                      [c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
    "hmx_hs_solve_f64": [vp, vp, c_int, c_int, c_int, vp] + [c_int, vp, vp, vp] * 4 +
                        [c_int, vp, vp, vp, vp],
}
_RESTYPE = {"hmx_last_error": c_char_p, "hmx_version": c_char_p}
_RESTYPE_ARGS = {"hmx_hs_forward_workspace": ([c_int, c_int, c_int], c_int64),
                 "hmx_hs_adjoint_workspace": ([c_int, c_int, c_int, c_int], c_int64)}

_lib = None
_load_error = None


def lib():
    """The loaded library (raises HindemithError if it cannot be loaded)."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if _load_error is not None:
        raise HindemithError(_load_error)
    if not os.path.exists(LIB_PATH):
        _load_error = (f"libhmx.so not found at {LIB_PATH}; build it with "
                       "`python -m hindemith_amd.build` (there is no CPU fallback)")
        raise HindemithError(_load_error)
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the host
        _load_error = f"could not load {LIB_PATH}: {e}"
        raise HindemithError(_load_error) from e
    for name, args in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = c_int
    for name, rt in _RESTYPE.items():
        fn = getattr(L, name)
        fn.argtypes = []
        fn.restype = rt
    for name, (args, rt) in _RESTYPE_ARGS.items():
        fn = getattr(L, name, None)
        if fn is not None:
            fn.argtypes = args
            fn.restype = rt
    _lib = L
    return L


def last_error():
    return lib().hmx_last_error().decode(errors="replace")


def check(rc, what=""):
    if rc != HMX_OK:
        msg = last_error()
        raise HindemithError(f"{what}: {_STATUS.get(rc, 'error')} ({rc}): {msg}", code=rc)
    return rc
