"""Device-tensor front doors to the hot-path kernels of libhmx.

Each function takes torch tensors that already live on the GPU (fp32,
contiguous), launches on torch's current stream and returns device tensors.
Stencils are ``(weights, offx, offy)`` triples of host arrays, the same data a
``hlib.Stencil`` carries.  These are what the autograd Function, the batched
solver and the benchmark call; the ``@hm`` front end goes through the record
API instead.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import HindemithError, check, lib

__all__ = [
    "taps_of", "stencil2d", "dia2d_spmv", "csr_spmv", "CSRBins", "csr_spmv_binned",
    "CSRRowBlocks", "csr_spmv_rowblocks", "CSRSlices", "csr_spmv_sliced",
    "sum2d", "hs_setup", "hs_jacobi", "hs_solve", "hs_tune", "hs_variants", "hs_config",
    "hs_recip_check", "stream_ptr",
    "hs_precondition", "hs_matrix_powers", "warp_img2d", "dia_spmv", "brox_set_matrix",
    "brox_set_redblack_matrix", "brox_precondition",
]


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _dev(t, name, dtype=torch.float32):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise HindemithError(f"{name} must be a CUDA/HIP tensor (there is no CPU path)")
    if t.dtype != dtype:
        raise HindemithError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise HindemithError(f"{name} must be contiguous")
    return t


class _Taps:
    """Host-side tap arrays kept alive for the duration of a call."""

    def __init__(self, st):
        if hasattr(st, "data") and hasattr(st, "offx"):
            w, ox, oy = st.data, st.offx, st.offy
        else:
            w, ox, oy = st
        self.w = np.ascontiguousarray(w, dtype=np.float32)
        self.ox = np.ascontiguousarray(ox, dtype=np.int32)
        self.oy = np.ascontiguousarray(oy, dtype=np.int32)
        if not (len(self.w) == len(self.ox) == len(self.oy)):
            raise HindemithError("stencil weights and offsets differ in length")

    def args(self):
        return [len(self.w), self.w.ctypes.data_as(ctypes.c_void_p),
                self.ox.ctypes.data_as(ctypes.c_void_p),
                self.oy.ctypes.data_as(ctypes.c_void_p)]


def taps_of(st):
    return _Taps(st)


def stencil2d(x, st, out=None):
    """a1: y = S * x (clamped borders)."""
    _dev(x, "x")
    H, W = x.shape
    y = torch.empty_like(x) if out is None else _dev(out, "out")
    t = _Taps(st)
    check(lib().hmx_stencil2d(_ptr(x), _ptr(y), H, W, *t.args(), stream_ptr()), "stencil2d")
    return y


def dia2d_spmv(A, offx, offy, x, out=None):
    """a2: y = A * x for a 2D DIA matrix A (nd x H x W planes)."""
    _dev(A, "A")
    _dev(x, "x")
    nd, H, W = A.shape
    if tuple(x.shape) != (H, W):
        raise HindemithError("dia2d_spmv: x shape does not match the matrix planes")
    ox = np.ascontiguousarray(offx, np.int32)
    oy = np.ascontiguousarray(offy, np.int32)
    y = torch.empty_like(x) if out is None else _dev(out, "out")
    check(lib().hmx_dia2d_spmv(_ptr(A), nd, ox.ctypes.data_as(ctypes.c_void_p),
                               oy.ctypes.data_as(ctypes.c_void_p), _ptr(x), _ptr(y), H, W,
                               stream_ptr()), "dia2d_spmv")
    return y


def csr_spmv(indptr, indices, data, x, n_cols=None, out=None, lanes_per_row=0):
    """a3: y = A * x, CSR with int32 indices (device tensors)."""
    _dev(indptr, "indptr", torch.int32)
    _dev(indices, "indices", torch.int32)
    _dev(data, "data")
    _dev(x, "x")
    M = indptr.numel() - 1
    y = torch.empty(M, device=x.device, dtype=torch.float32) if out is None else out
    check(lib().hmx_csr_spmv(M, x.numel() if n_cols is None else n_cols, data.numel(),
                             _ptr(indptr), _ptr(indices), _ptr(data), _ptr(x), _ptr(y),
                             int(lanes_per_row), stream_ptr()), "csr_spmv")
    return y


class CSRBins:
    """Row binning of a CSR matrix by row length (built once per matrix)."""

    def __init__(self, indptr_host, device):
        ip = np.ascontiguousarray(indptr_host, np.int32)
        M = len(ip) - 1
        bins = np.empty(max(M, 1), np.int32)
        counts = np.zeros(4, np.int32)
        check(lib().hmx_csr_bin_rows(M, ip.ctypes.data_as(ctypes.c_void_p),
                                     bins.ctypes.data_as(ctypes.c_void_p),
                                     counts.ctypes.data_as(ctypes.c_void_p)), "csr_bin_rows")
        self.M = M
        self.counts = counts
        self.bins = torch.from_numpy(bins[:max(M, 1)]).to(device)


class CSRRowBlocks:
    """Row blocks of a CSR matrix (<= 2048 entries, <= 256 rows each; built
    once per matrix on the host) for hmx_csr_spmv_rowblocks."""

    def __init__(self, indptr_host, device):
        ip = np.ascontiguousarray(indptr_host, np.int32)
        M = len(ip) - 1
        blocks = np.empty(M + 1, np.int32)
        nb = ctypes.c_int(0)
        check(lib().hmx_csr_row_blocks(M, ip.ctypes.data_as(ctypes.c_void_p),
                                       blocks.ctypes.data_as(ctypes.c_void_p),
                                       ctypes.byref(nb)), "csr_row_blocks")
        self.M = M
        self.nblocks = nb.value
        self.blocks = torch.from_numpy(blocks[:nb.value + 1].copy()).to(device)


def csr_spmv_rowblocks(indptr, indices, data, x, rb, out=None):
    """y = A * x with the row-block launch (rb: CSRRowBlocks of A)."""
    y = torch.empty(rb.M, device=x.device, dtype=torch.float32) if out is None else out
    check(lib().hmx_csr_spmv_rowblocks(rb.M, _ptr(indptr), _ptr(indices), _ptr(data), _ptr(x),
                                       _ptr(y), _ptr(rb.blocks), rb.nblocks, stream_ptr()),
          "csr_spmv_rowblocks")
    return y


class CSRSlices:
    """Column-sliced form of a CSR matrix (one column range per XCD; built once
    per sparsity pattern on the host) for hmx_csr_spmv_sliced.  ``values(data)``
    permutes a device value array of the original matrix into slice order."""

    def __init__(self, indptr_host, indices_host, n_cols, device):
        ip = np.ascontiguousarray(indptr_host, np.int32)
        ix = np.ascontiguousarray(indices_host, np.int32)
        M = len(ip) - 1
        nnz = int(ip[-1])
        L = lib()
        S = L.hmx_csr_slice_count()
        sptr = np.empty(S * (M + 1), np.int32)
        sidx = np.empty(max(nnz, 1), np.int32)
        perm = np.empty(max(nnz, 1), np.int32)
        check(L.hmx_csr_slice_build(M, int(n_cols), ip.ctypes.data_as(ctypes.c_void_p),
                                    ix.ctypes.data_as(ctypes.c_void_p),
                                    sptr.ctypes.data_as(ctypes.c_void_p),
                                    sidx.ctypes.data_as(ctypes.c_void_p),
                                    perm.ctypes.data_as(ctypes.c_void_p)), "csr_slice_build")
        self.M, self.nnz, self.slices = M, nnz, S
        self.sptr = torch.from_numpy(sptr).to(device)
        self.sidx = torch.from_numpy(sidx).to(device)
        self.perm = torch.from_numpy(perm).to(device)
        self.partial = torch.empty(max(S * M, 1), device=device, dtype=torch.float32)

    def values(self, data, out=None):
        _dev(data, "data")
        sd = torch.empty(max(self.nnz, 1), device=data.device, dtype=torch.float32) \
            if out is None else out
        check(lib().hmx_csr_slice_values(self.nnz, _ptr(self.perm), _ptr(data), _ptr(sd),
                                         stream_ptr()), "csr_slice_values")
        return sd


def csr_spmv_sliced(sl, sdata, x, out=None):
    """y = A * x from the column-sliced form (sl: CSRSlices of A, sdata:
    sl.values(data))."""
    _dev(x, "x")
    y = torch.empty(sl.M, device=x.device, dtype=torch.float32) if out is None else out
    check(lib().hmx_csr_spmv_sliced(sl.M, _ptr(sl.sptr), _ptr(sl.sidx), _ptr(sdata), _ptr(x),
                                    _ptr(sl.partial), _ptr(y), stream_ptr()), "csr_spmv_sliced")
    return y


def csr_spmv_binned(indptr, indices, data, x, bins, out=None):
    M = indptr.numel() - 1
    y = torch.empty(M, device=x.device, dtype=torch.float32) if out is None else out
    check(lib().hmx_csr_spmv_binned(M, _ptr(indptr), _ptr(indices), _ptr(data), _ptr(x),
                                    _ptr(y), _ptr(bins.bins),
                                    bins.counts.ctypes.data_as(ctypes.c_void_p), stream_ptr()),
          "csr_spmv_binned")
    return y


_SUM_PARTIALS = 1024


def sum2d(x, partials=None, out=None):
    """a7: deterministic sum of all elements -> 1-element fp32 device tensor."""
    _dev(x, "x")
    if partials is None:
        partials = torch.empty(_SUM_PARTIALS, device=x.device, dtype=torch.float64)
    y = torch.empty(1, device=x.device, dtype=torch.float32) if out is None else out
    check(lib().hmx_sum2d(_ptr(x), x.numel(), _ptr(partials), _ptr(y), stream_ptr()), "sum2d")
    return y


def hs_setup(I1, I2, Dx, Dy, Dt, out=None):
    """a10: (Ix, Iy, It) of the canonical program."""
    _dev(I1, "I1")
    _dev(I2, "I2")
    H, W = I1.shape
    if out is None:
        out = torch.empty((3, H, W), device=I1.device, dtype=torch.float32)
    Ix, Iy, It = out[0], out[1], out[2]
    tx, ty, tt = _Taps(Dx), _Taps(Dy), _Taps(Dt)
    check(lib().hmx_hs_setup(_ptr(I1), _ptr(I2), H, W, *tx.args(), *ty.args(), *tt.args(),
                             _ptr(Ix), _ptr(Iy), _ptr(It), stream_ptr()), "hs_setup")
    return Ix, Iy, It


def hs_jacobi(Ix, Iy, It, alpha2, Avg, iters, u, v, work=None, den=None, variant=None, k=None):
    """a9: ``iters`` fused Jacobi iterations, (u, v) updated in place."""
    for n, t in (("Ix", Ix), ("Iy", Iy), ("It", It), ("u", u), ("v", v)):
        _dev(t, n)
    H, W = Ix.shape
    if work is None:
        work = torch.empty((2, H, W), device=Ix.device, dtype=torch.float32)
    ta = _Taps(Avg)
    dptr = _ptr(den) if den is not None else ctypes.c_void_p(0)
    common = [_ptr(Ix), _ptr(Iy), _ptr(It), dptr, ctypes.c_float(alpha2), H, W, *ta.args(),
              int(iters), _ptr(u), _ptr(v), _ptr(work[0]), _ptr(work[1])]
    if variant is None:
        check(lib().hmx_hs_jacobi(*common, stream_ptr()), "hs_jacobi")
    else:
        check(lib().hmx_hs_jacobi_cfg(*common, int(variant), int(k), stream_ptr()), "hs_jacobi")
    return u, v


def hs_tune(Ix, Iy, It, alpha2, Avg):
    """Autotune (variant, k) for this image size; buffers are scratch copies."""
    H, W = Ix.shape
    u = torch.zeros_like(Ix)
    v = torch.zeros_like(Ix)
    work = torch.empty((2, H, W), device=Ix.device, dtype=torch.float32)
    ta = _Taps(Avg)
    vo, ko = ctypes.c_int(0), ctypes.c_int(0)
    check(lib().hmx_hs_tune(_ptr(Ix), _ptr(Iy), _ptr(It), ctypes.c_float(alpha2), H, W,
                            *ta.args(), _ptr(u), _ptr(v), _ptr(work[0]), _ptr(work[1]),
                            stream_ptr(), ctypes.byref(vo), ctypes.byref(ko)), "hs_tune")
    return vo.value, ko.value


def hs_config(H, W):
    """(variant, k) the fused Jacobi launches use for an H x W image."""
    vo, ko = ctypes.c_int(), ctypes.c_int()
    check(lib().hmx_hs_config(int(H), int(W), ctypes.byref(vo), ctypes.byref(ko)), "hs_config")
    return vo.value, ko.value


def hs_recip_check(e_lo=-24, e_hi=24):
    """Number of b = +-(1.m) * 2^e, e_lo <= e <= e_hi, for which the device
    reciprocal of the reciprocal quotient differs from the correctly rounded
    1/b (NaN outside [2^-24, 2^24]).  Zero means the quotient's y = RN(1/b)."""
    bad = torch.zeros(1, dtype=torch.int64, device="cuda")
    check(lib().hmx_hs_recip_check(int(e_lo), int(e_hi), ctypes.c_void_p(bad.data_ptr()),
                                   stream_ptr()), "hs_recip_check")
    return int(bad.item())


def hs_variants():
    L = lib()
    out = []
    for i in range(L.hmx_hs_variant_count()):
        tw, th, nt, km = (ctypes.c_int() for _ in range(4))
        check(L.hmx_hs_variant_info(i, ctypes.byref(tw), ctypes.byref(th), ctypes.byref(nt),
                                    ctypes.byref(km)), "variant_info")
        strip = L.hmx_hs_variant_strip(i)
        out.append(dict(tile_w=tw.value, tile_h=th.value, threads=nt.value, kmax=km.value,
                        strip=strip,
                        kernel="hs_strip_kernel" if strip > 0 else "hs_tb_kernel"))
    return out


def hs_solve(I1, I2, alpha2, taps, iters, u=None, v=None, workspace=None):
    """a9+a10: full solve from (u, v) (zeros by default); returns (u, v)."""
    _dev(I1, "I1")
    _dev(I2, "I2")
    H, W = I1.shape
    u = torch.zeros_like(I1) if u is None else _dev(u, "u")
    v = torch.zeros_like(I1) if v is None else _dev(v, "v")
    if workspace is None:
        workspace = torch.empty((5, H, W), device=I1.device, dtype=torch.float32)
    ts = [_Taps(taps[k]) for k in ("Dx", "Dy", "Dt", "Avg")]
    args = []
    for t in ts:
        args += t.args()
    check(lib().hmx_hs_solve(_ptr(I1), _ptr(I2), H, W, ctypes.c_float(alpha2), *args,
                             int(iters), _ptr(u), _ptr(v), _ptr(workspace), stream_ptr()),
          "hs_solve")
    return u, v


def hs_precondition(D, diag0, diag1, offdiag, x0, x1, out=None):
    """a15: (y0, y1) = M^-1 (x0, x1), the 2x2 block-Jacobi step of the
    reference's applyHSPreconditioner(D, diag0, diag1, offdiag, x0, x1); D is the
    stencil whose weight 2 enters the diagonal blocks."""
    for n, t in (("diag0", diag0), ("diag1", diag1), ("offdiag", offdiag), ("x0", x0),
                 ("x1", x1)):
        _dev(t, n)
    H, W = x0.shape
    for t in (diag0, diag1, offdiag, x1):
        if tuple(t.shape) != (H, W):
            raise HindemithError("hs_precondition: all planes must have the same shape")
    taps = _Taps(D)
    y0, y1 = (torch.empty_like(x0), torch.empty_like(x0)) if out is None else out
    check(lib().hmx_hs_precondition(_ptr(diag0), _ptr(diag1), _ptr(offdiag), _ptr(x0), _ptr(x1),
                                    _ptr(_dev(y0, "y0")), _ptr(_dev(y1, "y1")), H, W,
                                    len(taps.w), taps.w.ctypes.data_as(ctypes.c_void_p),
                                    stream_ptr()), "hs_precondition")
    return y0, y1


def hs_matrix_powers(lam, Ix, Iy, r0, r1, steps, out=None):
    """a16: V[0] = (r0, r1), V[s] = A V[s-1] for s < steps (the reference's
    hs_matrix_powers(lam, Ix, Iy, r0, r1) with defines['STEPS'] = steps);
    returns (V0, V1), each steps x H x W."""
    for n, t in (("Ix", Ix), ("Iy", Iy), ("r0", r0), ("r1", r1)):
        _dev(t, n)
    H, W = r0.shape
    for t in (Ix, Iy, r1):
        if tuple(t.shape) != (H, W):
            raise HindemithError("hs_matrix_powers: all planes must have the same shape")
    if out is None:
        V0 = torch.empty((steps, H, W), device=r0.device, dtype=torch.float32)
        V1 = torch.empty_like(V0)
    else:
        V0, V1 = (_dev(t, "V") for t in out)
    check(lib().hmx_hs_matrix_powers(ctypes.c_float(lam), _ptr(Ix), _ptr(Iy), _ptr(r0),
                                     _ptr(r1), _ptr(V0), _ptr(V1), H, W, int(steps),
                                     stream_ptr()), "hs_matrix_powers")
    return V0, V1


def warp_img2d(src, offx, offy, out=None):
    """a17: src bilinearly warped by the offset field (offx, offy), the
    reference's warp_img2d(src, offx, offy)."""
    for n, t in (("src", src), ("offx", offx), ("offy", offy)):
        _dev(t, n)
    H, W = src.shape
    if tuple(offx.shape) != (H, W) or tuple(offy.shape) != (H, W):
        raise HindemithError("warp_img2d: offsets must have the image's shape")
    dst = torch.empty_like(src) if out is None else _dev(out, "out")
    check(lib().hmx_warp_img2d(_ptr(src), _ptr(offx), _ptr(offy), _ptr(dst), H, W,
                               stream_ptr()), "warp_img2d")
    return dst


def dia_spmv(data, offsets, x, out=None):
    """1-D DIA SpMV of a square scipy-style dia_matrix: data (ndiags x n) on the
    device, offsets on the host; y[i] = sum_d data[d][i+o_d] * x[i+o_d]."""
    _dev(data, "data")
    _dev(x, "x")
    nd, n = data.shape
    if x.numel() != n:
        raise HindemithError("dia_spmv: vector length differs from the matrix size")
    off = np.ascontiguousarray(offsets, np.int32)
    if len(off) != nd:
        raise HindemithError("dia_spmv: one offset per diagonal")
    y = torch.empty_like(x) if out is None else _dev(out, "out")
    check(lib().hmx_dia_spmv(_ptr(data), nd, off.ctypes.data_as(ctypes.c_void_p), _ptr(x), _ptr(y),
                             n, stream_ptr()), "dia_spmv")
    return y


def brox_set_matrix(A, psi, alpha):
    """a19: planes 0..4 of the SpMat2D planes A (nd x H x W, in place) become the
    psi-weighted Laplacian (the reference's set_brox_matrix)."""
    _dev(A, "A")
    _dev(psi, "psi")
    nd, H, W = A.shape
    if tuple(psi.shape) != (H, W):
        raise HindemithError("brox_set_matrix: psi must be shaped like the planes")
    check(lib().hmx_brox_set_matrix(_ptr(A), nd, _ptr(psi), ctypes.c_float(alpha), H, W,
                                    stream_ptr()), "brox_set_matrix")
    return A


def brox_set_redblack_matrix(A, psi, alpha, out=None):
    """a19: off-diagonals into planes 0..3 of A (in place); returns the diagonal."""
    _dev(A, "A")
    _dev(psi, "psi")
    nd, H, W = A.shape
    if tuple(psi.shape) != (H, W):
        raise HindemithError("brox_set_redblack_matrix: psi must be shaped like the planes")
    diag = torch.empty_like(psi) if out is None else _dev(out, "out")
    check(lib().hmx_brox_set_redblack_matrix(_ptr(diag), _ptr(A), nd, _ptr(psi),
                                             ctypes.c_float(alpha), H, W, stream_ptr()),
          "brox_set_redblack_matrix")
    return diag


def brox_precondition(A, diag0, diag1, offdiag, x0, x1, out=None):
    """a19: the reference's applyBroxPreconditioner(A, diag0, diag1, offdiag, x0, x1)."""
    for n, t in (("A", A), ("diag0", diag0), ("diag1", diag1), ("offdiag", offdiag),
                 ("x0", x0), ("x1", x1)):
        _dev(t, n)
    nd, H, W = A.shape
    for t in (diag0, diag1, offdiag, x0, x1):
        if tuple(t.shape) != (H, W):
            raise HindemithError("brox_precondition: planes must be shaped like the matrix")
    y0, y1 = (torch.empty_like(x0), torch.empty_like(x0)) if out is None else out
    check(lib().hmx_brox_precondition(_ptr(A), nd, _ptr(diag0), _ptr(diag1), _ptr(offdiag),
                                      _ptr(x0), _ptr(x1), _ptr(_dev(y0, "y0")),
                                      _ptr(_dev(y1, "y1")), H, W, stream_ptr()),
          "brox_precondition")
    return y0, y1
