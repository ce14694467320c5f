"""Operand types of the Hindemith front end and their NumPy meaning.

Mirrors the reference's ``hlib.py`` for the Horn-Schunck path: ``Stencil``
(hlib.py:166-191), ``SpMat2D`` (hlib.py:140-164) and ``sum2d`` (hlib.py:83-84).
Without ``@hm`` a program runs on these NumPy definitions.  They follow the
generated kernels' semantics, which the reference's NumPy twin does not in two
places (both documented divergences of the reference itself):

* ``Stencil * x`` clamps each coordinate (Stencil2DMVM.cpp:42-45).  The
  reference's twin pads with ``mode='edge'`` too but mis-sizes its slices for
  offsets that are not symmetric about 0 (hlib.py:190), so the 2x2 derivative
  stencils of Horn-Schunck raise there; here they work.
* ``SpMat2D * x`` reads the coefficient at the neighbour index p+o, as the
  kernel does (SpMat2D.cpp:211-216); the reference's twin reads it at p
  (hlib.py:163).  Both agree for planes that are constant per diagonal.
"""
from __future__ import annotations

import numpy as np

__all__ = ["Stencil", "SpMat2D", "sum2d"]


class Stencil:
    """Constant-coefficient stencil: ``data[d]`` weight at offset
    ``(offx[d], offy[d])`` (column, row)."""

    def __init__(self, data, offx, offy):
        self.data = np.asarray(data)
        self.offx = np.asarray(offx)
        self.offy = np.asarray(offy)
        self.dtype = self.data.dtype
        if not (self.data.shape == self.offx.shape == self.offy.shape and self.data.ndim == 1):
            raise ValueError("Stencil data, offx and offy must be 1-D arrays of equal length")

    def __mul__(self, x):
        if not isinstance(x, np.ndarray) or x.ndim != 2:
            return NotImplemented
        H, W = x.shape
        rows = np.arange(H)
        cols = np.arange(W)
        y = np.zeros((H, W), dtype=np.result_type(x.dtype, self.data.dtype))
        for d in range(len(self.data)):
            r = np.clip(rows + int(self.offy[d]), 0, H - 1)
            c = np.clip(cols + int(self.offx[d]), 0, W - 1)
            y += self.data[d] * x[np.ix_(r, c)]
        return y.astype(x.dtype, copy=False)


class SpMat2D:
    """2D diagonal matrix: ``data`` is ndiags x H x W, one coefficient plane per
    diagonal at offset ``(offx[d], offy[d])``."""

    def __init__(self, data, offx, offy):
        self.data = np.asarray(data)
        self.ndiags, self.dim0, self.dim1 = self.data.shape
        self.dtype = self.data.dtype
        self.offx = np.asarray(offx)
        self.offy = np.asarray(offy)

    def __mul__(self, x):
        if not isinstance(x, np.ndarray) or x.ndim != 2:
            return NotImplemented
        H, W = x.shape
        y = np.zeros((H, W), dtype=np.result_type(x.dtype, self.data.dtype))
        for d in range(self.ndiags):
            ox, oy = int(self.offx[d]), int(self.offy[d])
            r0, r1 = max(0, -oy), min(H, H - oy)
            c0, c1 = max(0, -ox), min(W, W - ox)
            if r0 >= r1 or c0 >= c1:
                continue
            nb = (slice(r0 + oy, r1 + oy), slice(c0 + ox, c1 + ox))
            y[r0:r1, c0:c1] += self.data[d][nb] * x[nb]
        return y.astype(x.dtype, copy=False)


def sum2d(x):
    """Sum of all elements (hlib.py:83-84)."""
    return np.sum(np.asarray(x).flatten())
