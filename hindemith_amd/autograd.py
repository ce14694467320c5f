"""Differentiable Horn-Schunck solve (torch.autograd) on the gfx950 kernels.

``horn_schunck(I1, I2, alpha, iters=200, ckpt_every=20)`` takes B x H x W
frame stacks and per-pair ``alpha`` (B,) on the GPU and returns the flow
``(u, v)``.  The forward pass runs the fused, temporally blocked Jacobi kernel
and keeps the state every ``ckpt_every`` iterations; the backward pass
(hand-written adjoint kernel, libhmx ``hmx_hs_adjoint``) recomputes one
segment at a time and walks it backwards, so memory is
(iters/ckpt_every + ckpt_every) state pairs instead of iters.  The forward
keeps the last segment's linearisation (up to ckpt_every more pairs of
planes), so the backward recomputes every segment but that one.
Gradients flow to I1, I2, alpha and, when given, the initial flow u0, v0.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import HindemithError, check, lib
from .kernels import _Taps, _dev, _ptr, stream_ptr

__all__ = ["HornSchunckFunction", "horn_schunck", "default_taps", "hs_solve_f64"]


def default_taps():
    from .synthetic import hs_stencil_taps
    return hs_stencil_taps()


def _tap_args(taps):
    ts = [_Taps(taps[k]) for k in ("Dx", "Dy", "Dt", "Avg")]
    args = []
    for t in ts:
        args += t.args()
    return ts, args


class HornSchunckFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, I1, I2, alpha, u0, v0, iters, ckpt_every, taps):
        for n, t in (("I1", I1), ("I2", I2), ("alpha", alpha)):
            _dev(t, n)
        B, H, W = I1.shape
        if I2.shape != I1.shape or alpha.shape != (B,):
            raise HindemithError("horn_schunck: I1, I2 must be B x H x W and alpha (B,)")
        u = torch.zeros_like(I1) if u0 is None else _dev(u0, "u0").clone()
        v = torch.zeros_like(I1) if v0 is None else _dev(v0, "v0").clone()
        L = lib()
        nck = L.hmx_hs_ckpt_count(int(iters), int(ckpt_every))
        ckpt = torch.empty((max(nck, 1), 2, B, H, W), device=I1.device, dtype=torch.float32)
        ws = torch.empty(int(L.hmx_hs_forward_workspace(B, H, W)) // 4 + 1, device=I1.device,
                         dtype=torch.float32)
        keep, targs = _tap_args(taps)
        # With a gradient to come, the forward also keeps its last checkpoint
        # segment's linearisation (the backward then recomputes one segment
        # fewer) and its derivative planes (in ws); the gradients are the same
        # bits as without them.
        want = any(ctx.needs_input_grad[:5]) and int(iters) > 0
        last = int(iters) - (nck - 1) * int(ckpt_every) if want else 0
        lin = (torch.empty((2 * last, B, H, W), device=I1.device, dtype=torch.float32)
               if want else None)
        check(L.hmx_hs_forward_ckpt(_ptr(I1), _ptr(I2), B, H, W, _ptr(alpha), *targs,
                                    int(iters), int(ckpt_every), _ptr(u), _ptr(v), _ptr(ckpt),
                                    _ptr(lin) if lin is not None else ctypes.c_void_p(0),
                                    _ptr(ws), stream_ptr()), "hs_forward_ckpt")
        if lin is None:
            lin = torch.empty(0, device=I1.device)
        ctx.save_for_backward(I1, I2, alpha, ckpt, lin, ws)
        ctx.iters, ctx.ckpt_every, ctx.taps = int(iters), int(ckpt_every), taps
        ctx.has_u0 = u0 is not None
        ctx.kept = want
        return u, v

    @staticmethod
    def backward(ctx, gu, gv):
        I1, I2, alpha, ckpt, lin, fws = ctx.saved_tensors
        B, H, W = I1.shape
        gu = torch.zeros_like(I1) if gu is None else gu.contiguous()
        gv = torch.zeros_like(I1) if gv is None else gv.contiguous()
        L = lib()
        gI1 = torch.empty_like(I1)
        gI2 = torch.empty_like(I1)
        ga = torch.empty_like(alpha)
        gu0 = torch.empty_like(I1) if ctx.has_u0 else None
        gv0 = torch.empty_like(I1) if ctx.has_u0 else None
        ws = torch.empty(int(L.hmx_hs_adjoint_workspace(B, H, W, ctx.ckpt_every)) // 4 + 1,
                         device=I1.device, dtype=torch.float32)
        keep, targs = _tap_args(ctx.taps)
        null = ctypes.c_void_p(0)
        check(L.hmx_hs_adjoint(_ptr(I1), _ptr(I2), B, H, W, _ptr(alpha), *targs, ctx.iters,
                               ctx.ckpt_every, _ptr(ckpt), _ptr(lin) if ctx.kept else null,
                               _ptr(fws) if ctx.kept else null, _ptr(gu), _ptr(gv), _ptr(gI1),
                               _ptr(gI2), _ptr(ga), _ptr(gu0) if gu0 is not None else null,
                               _ptr(gv0) if gv0 is not None else null, _ptr(ws), stream_ptr()),
              "hs_adjoint")
        return gI1, gI2, ga, gu0, gv0, None, None, None


def horn_schunck(I1, I2, alpha, iters=200, ckpt_every=20, taps=None, u0=None, v0=None):
    """Differentiable HS flow of B pairs; returns (u, v), each B x H x W."""
    squeeze = I1.dim() == 2
    if squeeze:
        I1, I2 = I1[None], I2[None]
        u0 = None if u0 is None else u0[None]
        v0 = None if v0 is None else v0[None]
    if not torch.is_tensor(alpha):
        alpha = torch.full((I1.shape[0],), float(alpha), device=I1.device, dtype=torch.float32)
    alpha = alpha.reshape(-1)
    u, v = HornSchunckFunction.apply(I1.contiguous(), I2.contiguous(), alpha.contiguous(), u0,
                                     v0, iters, ckpt_every, taps or default_taps())
    return (u[0], v[0]) if squeeze else (u, v)


def hs_solve_f64(I1, I2, alpha, iters, taps=None):
    """fp64 forward solve of B pairs on the GPU (finite-difference checks)."""
    taps = taps or default_taps()
    B, H, W = I1.shape
    for n, t in (("I1", I1), ("I2", I2), ("alpha", alpha)):
        _dev(t, n, torch.float64)
    u = torch.zeros_like(I1)
    v = torch.zeros_like(I1)
    ws = torch.empty((5, B, H, W), device=I1.device, dtype=torch.float64)
    keep, targs = _tap_args(taps)
    check(lib().hmx_hs_solve_f64(_ptr(I1), _ptr(I2), B, H, W, _ptr(alpha), *targs, int(iters),
                                 _ptr(u), _ptr(v), _ptr(ws), stream_ptr()), "hs_solve_f64")
    return u, v
