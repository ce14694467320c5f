"""The C-ABI library loads and exports every function include/hmx.h declares."""
import ctypes
import os
import re

import pytest

from hindemith_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    hdr = open(os.path.join(ROOT, "include", "hmx.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|int64_t|const char \*|void)\s*\*?\s*(hmx_\w+)\s*\(",
                                 hdr)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("hmx_init", "hmx_compile", "hmx_bind", "hmx_run", "hmx_result_read",
                 "hmx_release", "hmx_stencil2d", "hmx_dia2d_spmv", "hmx_csr_spmv", "hmx_sum2d",
                 "hmx_hs_setup", "hmx_hs_jacobi", "hmx_hs_solve"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_no_gpu_needed_for_error_paths():
    L = _lib.lib()
    assert L.hmx_version().startswith(b"hmx")
    # invalid arguments are rejected before any HIP call
    rc = L.hmx_stencil2d(None, None, 4, 4, 1, None, None, None, None)
    assert rc == 1 and b"null" in L.hmx_last_error()
    rc = L.hmx_hs_jacobi_cfg(None, None, None, None, ctypes.c_float(1.0), 4, 4, 0, None, None,
                             None, 1, None, None, None, None, 0, 1, None)
    assert rc == 1
    assert L.hmx_run(123456) == 4  # unknown record


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "_load_error", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libhmx.so")
    with pytest.raises(_lib.HindemithError, match="no CPU fallback"):
        _lib.lib()
