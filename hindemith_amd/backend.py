"""Drop-in for the reference's ``backend`` extension module (backend.cpp).

Same three functions, same arguments and results, implemented over libhmx's
record C ABI instead of the Python-2 C extension + OpenCL code generator:

* ``xmlcompile(record_id, symtab_xml, sm_xml, defines_xml)`` -- backend.cpp:70-159
* ``run(record_id, kwargs) -> (retval, {})``                  -- backend.cpp:320-638
* ``free_ocl()``                                               -- backend.cpp:646-651

The reference's PyAST.py imports this module by name (PyAST.py:8) and calls
exactly these (PyAST.py:194, :198-200, :202-203).  Behaviour kept: kwargs are
bound by symbol name, a scalar result comes back as a Python number, one
result is returned bare and several as a tuple, and a returned kwarg array is
the caller's own array updated in place (the reference wraps the caller's
buffer, backend.cpp:186-276).  Behaviour changed, each a failure the reference
reported by printing and continuing or by crashing: a missing kwarg, a dtype
other than float32 for arrays, a non-contiguous array or an unknown record
raise ``HindemithError``; other results are fresh arrays the caller owns (the
reference returned views of backend memory that the next call overwrote).
"""
from __future__ import annotations

import ctypes
import threading
import xml.etree.ElementTree as ET

import numpy as np

from ._lib import HindemithError, check, lib

__all__ = ["xmlcompile", "run", "free_ocl", "init", "plan", "program_source"]

_lock = threading.Lock()
_symtabs = {}    # record id -> {name: (tag, attrs)}
_device = [None]

_DTYPES = {0: np.float32, 1: np.float64, 2: np.int64, 3: np.int32}


def init(device):
    """initbackend -> initialize_ocl (backend.cpp:669-675): select the GPU
    (-1: plan-only, no GPU needed)."""
    with _lock:
        if _device[0] != device:
            check(lib().hmx_init(int(device)), "init")
            _device[0] = device


def xmlcompile(record_id, symtab_xml, sm_xml, defines_xml):
    """Parse and plan a program record; allocates its device buffers."""
    if _device[0] is None:
        raise HindemithError("backend not initialised: select a device first")
    st = ET.fromstring(symtab_xml)
    syms = {e.get("name"): (e.tag, dict(e.attrib)) for e in st}
    check(lib().hmx_compile(int(record_id), symtab_xml.encode(), sm_xml.encode(),
                            defines_xml.encode()), "xmlcompile")
    with _lock:
        _symtabs[int(record_id)] = syms
    return None


def plan(record_id):
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().hmx_plan_text(int(record_id), buf, len(buf)), "plan")
    return buf.value.decode()


def program_source(record_id, compile=False):
    """HIP source of the kernels generated for the record's fused groups;
    with compile=True each is also built with hipRTC (raises on failure)."""
    buf = ctypes.create_string_buffer(1 << 20)
    check(lib().hmx_program_source(int(record_id), 1 if compile else 0, buf, len(buf)),
          "program_source")
    return buf.value.decode()


def _bind(rid, name, comp, arr, keep):
    keep.append(arr)
    check(lib().hmx_bind(rid, name.encode(), comp, arr.ctypes.data_as(ctypes.c_void_p),
                         arr.nbytes), f"bind {name}")


def _bind_symbol(rid, name, tag, attrs, obj, keep):
    if tag in ("Array", "Array2D"):
        if not isinstance(obj, np.ndarray):
            raise HindemithError(f"kwarg {name}: expected an ndarray")
        if obj.dtype != np.float32:
            raise HindemithError(f"kwarg {name}: arrays must be float32 on this backend, "
                                 f"got {obj.dtype}")
        if not (obj.flags["C_CONTIGUOUS"] or obj.flags["F_CONTIGUOUS"]):
            raise HindemithError(f"kwarg {name}: array must be contiguous")
        _bind(rid, name, 0, obj, keep)
    elif tag == "Scalar":
        _bind(rid, name, 0, np.array([obj], dtype=attrs.get("dtype", "float32")), keep)
    elif tag in ("Stencil", "SpMat2D"):
        if tag == "SpMat2D" and np.asarray(obj.data).dtype != np.float32:
            raise HindemithError(f"kwarg {name}: SpMat2D planes must be float32")
        # stencil weights are cast to float32; offsets to int32 (the kernels read
        # them as int, Stencil.cpp:110)
        _bind(rid, name, 0, np.ascontiguousarray(obj.data, dtype=np.float32), keep)
        _bind(rid, name, 1, np.ascontiguousarray(obj.offx, dtype=np.int32), keep)
        _bind(rid, name, 2, np.ascontiguousarray(obj.offy, dtype=np.int32), keep)
    elif tag == "SpMat":  # scipy dia_matrix: data (ndiags x n), offsets
        if obj.dtype != np.float32:
            raise HindemithError(f"kwarg {name}: dia_matrix data must be float32")
        _bind(rid, name, 0, np.ascontiguousarray(obj.data), keep)
        _bind(rid, name, 1, np.ascontiguousarray(obj.offsets, dtype=np.int32), keep)
    elif tag == "CSRMat":
        if obj.dtype != np.float32:
            raise HindemithError(f"kwarg {name}: csr_matrix data must be float32")
        idx, ptr = np.asarray(obj.indices), np.asarray(obj.indptr)
        lim = np.iinfo(np.int32).max
        if (idx.size and int(idx.max()) > lim) or int(ptr[-1]) > lim:
            raise HindemithError(f"kwarg {name}: indices exceed int32 (CSRMat.cpp:26-27)")
        _bind(rid, name, 0, np.ascontiguousarray(obj.data), keep)
        _bind(rid, name, 1, np.ascontiguousarray(idx, dtype=np.int32), keep)
        _bind(rid, name, 2, np.ascontiguousarray(ptr, dtype=np.int32), keep)
    else:
        raise HindemithError(f"kwarg {name}: symbol type {tag} is not supported")


def run(record_id, kwargs):
    """Bind kwargs, execute, and return (retval, {}) like backend.run."""
    rid = int(record_id)
    with _lock:
        syms = _symtabs.get(rid)
    if syms is None:
        raise HindemithError(f"no compiled record {rid}")
    keep = []
    for name, (tag, attrs) in syms.items():
        if attrs.get("source") != "Header":
            continue
        if name not in kwargs:
            raise HindemithError(f"kwarg {name} missing (the reference printed 'Variable data "
                                 "not found', backend.cpp:417-420)")
        _bind_symbol(rid, name, tag, attrs, kwargs[name], keep)
    check(lib().hmx_run(rid), "run")
    n = ctypes.c_int(0)
    check(lib().hmx_result_count(rid, ctypes.byref(n)), "result_count")
    outs = []
    for i in range(n.value):
        kind, dt, nd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        shape = (ctypes.c_int64 * 3)()
        nm = ctypes.c_char_p()
        check(lib().hmx_result_info(rid, i, ctypes.byref(kind), ctypes.byref(dt),
                                    ctypes.byref(nd), shape, ctypes.byref(nm)), "result_info")
        name = nm.value.decode()
        dtype = _DTYPES[dt.value]
        if kind.value == 3:  # scalar -> Python number (backend.cpp:277-306)
            val = np.zeros(1, dtype=dtype)
            check(lib().hmx_result_read(rid, i, val.ctypes.data_as(ctypes.c_void_p), val.nbytes),
                  "result_read")
            outs.append(int(val[0]) if dtype == np.int64 else float(val[0]))
            continue
        if nd.value == 3:  # Array3D (HSMatrixPowers result): C order
            shp, order = (int(shape[0]), int(shape[1]), int(shape[2])), "C"
        else:
            shp = (int(shape[0]), int(shape[1])) if nd.value == 2 else (int(shape[0]),)
            order = "F" if shape[2] else "C"
        src = kwargs.get(name) if syms[name][1].get("source") == "Header" else None
        out = src if isinstance(src, np.ndarray) else np.empty(shp, dtype=dtype, order=order)
        check(lib().hmx_result_read(rid, i, out.ctypes.data_as(ctypes.c_void_p), out.nbytes),
              "result_read")
        outs.append(out)
    retval = None if not outs else (outs[0] if len(outs) == 1 else tuple(outs))
    return retval, {}


def free_ocl():
    """Release every record (the reference's free_ocl was a no-op)."""
    with _lock:
        ids = list(_symtabs)
        _symtabs.clear()
    for rid in ids:
        lib().hmx_release(rid)
    return None
