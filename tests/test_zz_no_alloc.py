"""hmx_sum2d_noalloc: the allocation-free form of hmx_sum2d, called through the
C ABI.  Its pass-1 partials live in the library's per-device scratch array, so
it takes no workspace; its result must equal hmx_sum2d's bit for bit (same
kernels, same fixed grid and fold order) and lie within one fp32 rounding of
the exact sum, on consecutive calls whose second input is smaller (the
scratch keeps the previous call's partials, which must not leak in).
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

HMX_SUM_PARTIALS = 1024


def _lib():
    from hindemith_amd._lib import lib
    L = lib()
    fn = L.hmx_sum2d_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return L


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _sum_with_workspace(L, x):
    part = torch.empty(HMX_SUM_PARTIALS, device="cuda", dtype=torch.float64)
    out = torch.empty(1, device="cuda", dtype=torch.float32)
    rc = L.hmx_sum2d(ctypes.c_void_p(x.data_ptr()), x.numel(), ctypes.c_void_p(part.data_ptr()),
                     ctypes.c_void_p(out.data_ptr()), _stream())
    assert rc == 0, L.hmx_last_error()
    return out


def _sum_noalloc(L, x):
    out = torch.full((1,), float("nan"), device="cuda", dtype=torch.float32)
    ptr = x.data_ptr() if x.numel() else 0
    rc = L.hmx_sum2d_noalloc(ctypes.c_void_p(ptr), x.numel(), ctypes.c_void_p(out.data_ptr()),
                             _stream())
    assert rc == 0, L.hmx_last_error()
    return out


def test_noalloc_sum_equals_workspace_sum_on_shrinking_inputs():
    L = _lib()
    rng = np.random.default_rng(21)
    # a large input, then a smaller one, then one smaller than the partials
    # grid (most partials are zero), then an empty one
    for n in (3_000_001, 777_777, 1000, 0):
        xh = rng.normal(size=n).astype(np.float32)
        x = torch.from_numpy(xh).cuda()
        got = _sum_noalloc(L, x)
        ref = _sum_with_workspace(L, x)
        torch.cuda.synchronize()
        g, r = got.cpu().numpy(), ref.cpu().numpy()
        assert g.tobytes() == r.tobytes(), (n, g, r)
        exact = float(np.sum(xh.astype(np.float64)))
        assert abs(float(g[0]) - exact) <= abs(exact) * 2.0 ** -23 + 1e-30, (n, g, exact)


def test_noalloc_rejects_bad_arguments():
    L = _lib()
    x = torch.ones(10, device="cuda")
    rc = L.hmx_sum2d_noalloc(ctypes.c_void_p(x.data_ptr()), 10, None, _stream())
    assert rc != 0 and b"null out" in L.hmx_last_error()
    out = torch.empty(1, device="cuda")
    rc = L.hmx_sum2d_noalloc(None, 10, ctypes.c_void_p(out.data_ptr()), _stream())
    assert rc != 0 and b"bad input" in L.hmx_last_error()
