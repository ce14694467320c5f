"""Many Horn-Schunck pairs of uneven sizes in one batched call.

``BatchedHS(I1s, I2s, alphas, iters)`` keeps per-pair device buffers and a
ctypes array of ``hmx_pair`` descriptors; ``run()`` makes one call of
``hmx_hs_batched``, which does the derivative setup and all Jacobi
iterations of every pair over one tile list of all pairs, captured in a
hipGraph on first use and replayed afterwards.  The default (stepped) form is
one launch per phase of k iterations over every pair's tiles; the persistent
form is one launch with a grid barrier between phases.  ``tune()`` picks the
form, tile shape and iterations per phase for this list of sizes.
``solve_each()`` is the plain alternative: one ``hmx_hs_solve`` per pair.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import HindemithError, check, lib
from .kernels import _Taps, _dev, stream_ptr

__all__ = ["BatchedHS", "hmx_pair", "BATCH_GRAPH", "BATCH_SYNC"]

BATCH_GRAPH = 1
BATCH_SYNC = 2


class hmx_pair(ctypes.Structure):
    _fields_ = [("I1", ctypes.c_void_p), ("I2", ctypes.c_void_p), ("u", ctypes.c_void_p),
                ("v", ctypes.c_void_p), ("work", ctypes.c_void_p), ("H", ctypes.c_int),
                ("W", ctypes.c_int), ("alpha", ctypes.c_float)]


def _sig():
    L = lib()
    tap = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p] * 4
    L.hmx_hs_batched.argtypes = ([ctypes.c_void_p, ctypes.c_int] + tap +
                                 [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_longlong, ctypes.c_void_p])
    L.hmx_hs_batched.restype = ctypes.c_int
    L.hmx_hs_batched_tune.argtypes = ([ctypes.c_void_p, ctypes.c_int] + tap +
                                      [ctypes.c_int, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int)])
    L.hmx_hs_batched_tune.restype = ctypes.c_int
    return L


def batched_variants(ntaps=8):
    """Tile shape, threads, max k and residency of every batched variant."""
    L = lib()
    out = []
    for i in range(L.hmx_hs_batched_variant_count()):
        tw, th, nt, km, api, own, cus = (ctypes.c_int() for _ in range(7))
        check(L.hmx_hs_batched_variant_info(i, ctypes.byref(tw), ctypes.byref(th),
                                            ctypes.byref(nt), ctypes.byref(km)),
              "batched_variant_info")
        check(L.hmx_hs_batched_resident(i, ntaps, ctypes.byref(api), ctypes.byref(own),
                                        ctypes.byref(cus)), "batched_resident")
        out.append(dict(tile_w=tw.value, tile_h=th.value, threads=nt.value, kmax=km.value,
                        api_per_cu=api.value, own_per_cu=own.value, cus=cus.value))
    return out


class BatchedHS:
    def __init__(self, I1s, I2s, alphas, iters=200, taps=None):
        if taps is None:
            from .synthetic import hs_stencil_taps
            taps = hs_stencil_taps()
        if not (len(I1s) == len(I2s) == len(alphas)) or not I1s:
            raise HindemithError("BatchedHS: need equally many I1, I2 and alpha (>= 1)")
        self.iters = int(iters)
        self.I1s = [_dev(t, "I1") for t in I1s]
        self.I2s = [_dev(t, "I2") for t in I2s]
        self.alphas = [float(a) for a in alphas]
        self.u = [torch.zeros_like(t) for t in self.I1s]
        self.v = [torch.zeros_like(t) for t in self.I1s]
        self.work = [torch.empty((5,) + tuple(t.shape), device=t.device, dtype=torch.float32)
                     for t in self.I1s]
        n = len(self.I1s)
        self.desc = (hmx_pair * n)()
        for i in range(n):
            H, W = self.I1s[i].shape
            if tuple(self.I2s[i].shape) != (H, W):
                raise HindemithError(f"BatchedHS: pair {i} frames differ in shape")
            self.desc[i] = hmx_pair(self.I1s[i].data_ptr(), self.I2s[i].data_ptr(),
                                    self.u[i].data_ptr(), self.v[i].data_ptr(),
                                    self.work[i].data_ptr(), H, W, self.alphas[i])
        self.taps = [_Taps(taps[k]) for k in ("Dx", "Dy", "Dt", "Avg")]
        self.taps_dict = taps
        self.variant, self.k = -1, 0

    def _tap_args(self):
        out = []
        for t in self.taps:
            out += t.args()
        return out

    def reset(self):
        for t in self.u + self.v:
            t.zero_()

    def run(self, graph=True, sync=False, spin_limit=0, variant=None, k=None):
        """One batched call over all pairs; returns the list of (u, v).  A
        persistent variant always synchronises and raises if its grid barrier
        gave up; sync=True makes the stepped form synchronise too."""
        self.reset()
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_SYNC if sync else 0)
        L = _sig()
        v = self.variant if variant is None else variant
        kk = self.k if k is None else k
        check(L.hmx_hs_batched(ctypes.cast(self.desc, ctypes.c_void_p), len(self.desc),
                               *self._tap_args(), self.iters, int(v), int(kk), flags,
                               int(spin_limit), stream_ptr()), "hs_batched")
        return list(zip(self.u, self.v))

    def tune(self):
        L = _sig()
        vo, ko = ctypes.c_int(), ctypes.c_int()
        check(L.hmx_hs_batched_tune(ctypes.cast(self.desc, ctypes.c_void_p), len(self.desc),
                                    *self._tap_args(), self.iters, stream_ptr(),
                                    ctypes.byref(vo), ctypes.byref(ko)), "hs_batched_tune")
        self.variant, self.k = vo.value, ko.value
        return self.variant, self.k

    def solve_each(self, variant=None, k=None):
        """The plain path: one solve per pair (setup + fused Jacobi).  With
        variant/k given, the Jacobi launches use that kernel variant (variant 0,
        k = 1 is one launch per iteration, the reference's launch structure)."""
        from .kernels import hs_jacobi, hs_setup, hs_solve
        t = self.taps_dict
        out = []
        for i in range(len(self.I1s)):
            self.u[i].zero_()
            self.v[i].zero_()
            a = self.alphas[i]
            a2 = float(torch.tensor(a, dtype=torch.float32) * torch.tensor(a, dtype=torch.float32))
            if variant is None:
                hs_solve(self.I1s[i], self.I2s[i], a2, t, self.iters, self.u[i], self.v[i],
                         workspace=self.work[i])
            else:
                w = self.work[i]
                Ix, Iy, It = hs_setup(self.I1s[i], self.I2s[i], t["Dx"], t["Dy"], t["Dt"],
                                      out=w[0:3])
                hs_jacobi(Ix, Iy, It, a2, t["Avg"], self.iters, self.u[i], self.v[i],
                          work=w[3:5], variant=variant, k=k)
            out.append((self.u[i], self.v[i]))
        return out
