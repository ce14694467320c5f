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


def test_csr_row_blocks_host():
    """The host-side row blocking behind hmx_csr_spmv_rowblocks: blocks cover
    the rows in order, hold at most 2048 entries and 256 rows, and a longer
    row stands alone (no GPU involved)."""
    import numpy as np
    from hindemith_amd._lib import lib
    rng = np.random.default_rng(5)
    for lens in (np.array([], np.int64), np.array([3]), np.array([5000]),
                 rng.integers(0, 40, 10000), np.concatenate([rng.integers(0, 5, 600), [4000],
                                                             np.zeros(700, np.int64), [2049]])):
        M = len(lens)
        ip = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        blocks = np.empty(M + 1, np.int32)
        nb = ctypes.c_int(-1)
        assert lib().hmx_csr_row_blocks(M, ip.ctypes.data_as(ctypes.c_void_p),
                                        blocks.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(nb)) == 0
        b = blocks[:nb.value + 1]
        assert b[0] == 0 and b[-1] == M and np.all(np.diff(b) >= 1)
        for r0, r1 in zip(b[:-1], b[1:]):
            n = ip[r1] - ip[r0]
            assert r1 - r0 <= 256
            assert n <= 2048 or r1 - r0 == 1


def test_csr_slice_build_host():
    """The host-side column slicing behind hmx_csr_spmv_sliced: slice s holds
    exactly the entries whose column falls in its range, a row keeps its
    entry order inside each slice, perm points back at each entry, and the
    sliced product regrouped in fp64 equals the original one (no GPU)."""
    import numpy as np
    import scipy.sparse as sp
    from hindemith_amd import synthetic as S
    from hindemith_amd._lib import lib
    L = lib()
    ns = L.hmx_csr_slice_count()
    assert ns == 8
    rng = np.random.default_rng(9)
    cases = [(np.array([0], np.int32), np.array([], np.int32), 5),       # no rows
             (np.array([0, 0, 0], np.int32), np.array([], np.int32), 3),  # empty rows, N < 8
             (np.array([0, 3], np.int32), np.array([2, 0, 1], np.int32), 3)]
    ip, ix, _ = S.powerlaw_csr(3000, rng, n_cols=2501)
    cases.append((ip, ix, 2501))
    for ip, ix, N in cases:
        M = len(ip) - 1
        nnz = int(ip[-1])
        sptr = np.empty(ns * (M + 1), np.int32)
        sidx = np.empty(max(nnz, 1), np.int32)
        perm = np.empty(max(nnz, 1), np.int32)
        rc = L.hmx_csr_slice_build(M, N, ip.ctypes.data_as(ctypes.c_void_p),
                                   ix.ctypes.data_as(ctypes.c_void_p),
                                   sptr.ctypes.data_as(ctypes.c_void_p),
                                   sidx.ctypes.data_as(ctypes.c_void_p),
                                   perm.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0, L.hmx_last_error()
        sw = -(-N // ns)
        sp_ = sptr.reshape(ns, M + 1)
        # slice s ends where slice s + 1 starts: offsets never decrease
        assert sp_[0, 0] == 0 and sp_[-1, -1] == nnz and np.all(np.diff(sptr) >= 0)
        sidx, perm = sidx[:nnz], perm[:nnz]
        assert np.array_equal(np.sort(perm), np.arange(nnz))
        assert np.array_equal(ix[perm], sidx)
        for s in range(ns):
            for i in range(M):
                got = perm[sp_[s, i]:sp_[s, i + 1]]
                want = [j for j in range(ip[i], ip[i + 1]) if ix[j] // sw == s]
                assert list(got) == want
        data = rng.normal(size=nnz)
        x = rng.normal(size=N)
        y = np.zeros(M)
        for s in range(ns):
            for i in range(M):
                k = slice(sp_[s, i], sp_[s, i + 1])
                y[i] += np.dot(data[perm[k]], x[sidx[k]])
        ref = sp.csr_matrix((data, ix, ip), shape=(M, N)) @ x if nnz else np.zeros(M)
        assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)
    # a column index outside [0, N) is rejected on the host
    bad_ip, bad_ix = np.array([0, 1], np.int32), np.array([7], np.int32)
    rc = L.hmx_csr_slice_build(1, 5, bad_ip.ctypes.data_as(ctypes.c_void_p),
                               bad_ix.ctypes.data_as(ctypes.c_void_p),
                               np.empty(16, np.int32).ctypes.data_as(ctypes.c_void_p),
                               np.empty(1, np.int32).ctypes.data_as(ctypes.c_void_p),
                               np.empty(1, np.int32).ctypes.data_as(ctypes.c_void_p))
    assert rc == 1 and b"out of range" in L.hmx_last_error()
