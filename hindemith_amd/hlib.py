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
* ``applyHSPreconditioner`` passes a block through where |det| <= 1e-15, the
  kernel's threshold (ApplyHSPreconditioner.cpp:73); the reference's twin uses
  1e-4 (hlib.py:38).
* ``hs_matrix_powers`` has no NumPy twin in the reference; this one follows
  HSMatrixPowers.cpp:72-126 and takes the number of powers from
  ``defines['STEPS']`` as the kernel does.
* ``set_brox_matrix`` stores each coefficient at the neighbour's index and
  leaves entries no statement writes alone, as the kernel does
  (SetBroxMatrix.cpp:71-75); the reference's twin stores them at the pixel and
  zeroes the planes first (hlib.py:71-82), matching its own pixel-indexed
  SpMat2D product.
* ``warp_img2d`` splits the offset as the kernel does (truncate, one step down
  when negative, WarpImg2D.cpp:37-50) and copies the source pixel when a tap
  leaves the image; the reference's twin scales the fractions by 1 - sign
  (hlib.py:128-129) and tests other bounds.
"""
from __future__ import annotations

import numpy as np

__all__ = ["Stencil", "SpMat2D", "sum2d", "applyHSPreconditioner", "hs_matrix_powers",
           "warp_img2d", "set_brox_matrix", "set_brox_redblack_matrix",
           "applyBroxPreconditioner", "red", "black"]


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


def _neighbour_count(H, W):
    n = np.zeros((H, W), np.float32)
    n[:, :-1] += 1
    n[1:, :] += 1
    n[:, 1:] += 1
    n[:-1, :] += 1
    return n


def applyHSPreconditioner(D, diag0, diag1, offdiag, x0, x1):
    """2x2 block-Jacobi step (ApplyHSPreconditioner.cpp:46-97): solve
    [[a, b], [b, d]] y = x per pixel with a = diag0 + n*w2, d = diag1 + n*w2,
    b = offdiag, n the in-image 4-neighbours, w2 = D.data[2]; y = x where
    |det| <= 1e-15."""
    x0 = np.asarray(x0)
    H, W = x0.shape
    s2 = np.float32(D.data[2])
    n = _neighbour_count(H, W)
    a = (diag0 + n * s2).astype(x0.dtype)
    d = (diag1 + n * s2).astype(x0.dtype)
    b = np.asarray(offdiag)
    det = a * d - b * b
    ok = np.abs(det.astype(np.float64)) > 1e-15
    safe = np.where(ok, det, 1)
    y0 = np.where(ok, (x0 * d - b * x1) / safe, x0)
    y1 = np.where(ok, (-b * x0 + a * x1) / safe, x1)
    return y0.astype(x0.dtype), y1.astype(x0.dtype)


def hs_matrix_powers(lam, Ix, Iy, r0, r1, steps=None):
    """V[0] = (r0, r1), V[s] = A V[s-1] for s < steps (HSMatrixPowers.cpp:72-126):
    A = [Ix (Ix u + Iy v), Iy (Ix u + Iy v)] + lam * (4-neighbour Neumann
    Laplacian).  steps defaults to defines['STEPS'] of the hindemith module."""
    if steps is None:
        from .hindemith import defines
        steps = int(defines["STEPS"])
    r0 = np.asarray(r0)
    H, W = r0.shape
    V0 = np.empty((steps, H, W), r0.dtype)
    V1 = np.empty_like(V0)
    V0[0], V1[0] = r0, r1
    for s in range(1, steps):
        u, v = V0[s - 1], V1[s - 1]
        dot = Ix * u + Iy * v
        U, V = Ix * dot, Iy * dot
        for nb_u, nb_v, sl in ((u[:, :-1], v[:, :-1], np.s_[:, 1:]),
                               (u[:, 1:], v[:, 1:], np.s_[:, :-1]),
                               (u[:-1, :], v[:-1, :], np.s_[1:, :]),
                               (u[1:, :], v[1:, :], np.s_[:-1, :])):
            U[sl] += lam * (u[sl] - nb_u)
            V[sl] += lam * (v[sl] - nb_v)
        V0[s], V1[s] = U, V
    return V0, V1


def warp_img2d(src, offx, offy):
    """Bilinear warp (WarpImg2D.cpp:33-75): src sampled at (x + offx, y + offy)
    where the four taps are inside the image, else src(x, y)."""
    src = np.asarray(src)
    H, W = src.shape
    ox, oy = np.asarray(offx), np.asarray(offy)
    mx, my = np.trunc(ox).astype(np.int64), np.trunc(oy).astype(np.int64)
    xf, yf = ox - mx, oy - my
    neg_x, neg_y = ox < 0, oy < 0
    mx, my = mx - neg_x, my - neg_y
    xf, yf = np.where(neg_x, 1 + xf, xf), np.where(neg_y, 1 + yf, yf)
    rr, cc = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    sx, sy = cc + mx, rr + my
    ok = (sx >= 0) & (sx + 1 < W) & (sy >= 0) & (sy + 1 < H)
    sx, sy = np.where(ok, sx, 0), np.where(ok, sy, 0)
    sx1, sy1 = np.minimum(sx + 1, W - 1), np.minimum(sy + 1, H - 1)
    val = (src[sy, sx] * (1 - xf) * (1 - yf) + src[sy, sx1] * xf * (1 - yf) +
           src[sy1, sx] * (1 - xf) * yf + src[sy1, sx1] * xf * yf)
    return np.where(ok, val, src).astype(src.dtype)


def _brox_planes(A, psi, alpha, planes):
    """Write the psi-weighted Laplacian's off-diagonals into planes (top, left,
    right, bottom) of A.data, each at the neighbour's index; return the diagonal."""
    psi = np.asarray(psi)
    H, W = psi.shape
    d = A.data
    t, l, r, b = planes
    mid = np.zeros((H, W), np.float32)
    d[t, :-1, :] = -alpha * psi[:-1, :]          # pixel (y, x) with y > 0 writes at y - 1
    mid[1:, :] += alpha * psi[:-1, :]
    d[l, :, :-1] = -alpha * psi[:, :-1]
    mid[:, 1:] += alpha * psi[:, :-1]
    d[r, :, 1:] = -alpha * psi[:, :-1]
    mid[:, :-1] += alpha * psi[:, :-1]
    d[b, 1:, :] = -alpha * psi[:-1, :]
    mid[:-1, :] += alpha * psi[:-1, :]
    return mid


def set_brox_matrix(A, psi, alpha):
    """SetBroxMatrix.cpp:34-79: planes 0..4 of A (offsets (0,-1), (-1,0), (0,0),
    (1,0), (0,1)) become the psi-weighted 4-neighbour Laplacian."""
    A.data[2] = _brox_planes(A, psi, alpha, (0, 1, 3, 4))


def set_brox_redblack_matrix(A, psi, alpha):
    """SetBroxRedblackMatrix.cpp:37-81: off-diagonals into planes 0..3, the
    diagonal returned as an array."""
    return _brox_planes(A, psi, alpha, (0, 1, 2, 3))


def applyBroxPreconditioner(A, diag0, diag1, offdiag, x0, x1):
    """ApplyBroxPreconditioner.cpp:46-72: 2x2 solve with a = diag0 + A[2],
    d = diag1 + A[2], b = offdiag; x passed through where |det| <= 1e-4."""
    x0 = np.asarray(x0)
    a = diag0 + A.data[2]
    d = diag1 + A.data[2]
    b = np.asarray(offdiag)
    det = a * d - b * b
    ok = np.abs(det.astype(np.float64)) > 1e-4
    safe = np.where(ok, det, 1)
    y0 = np.where(ok, (x0 * d - b * x1) / safe, x0)
    y1 = np.where(ok, (-b * x0 + a * x1) / safe, x1)
    return y0.astype(x0.dtype), y1.astype(x0.dtype)


def red(x):
    """RedArray2D.cpp:20-24: x where (column + row) is even, else 0."""
    x = np.asarray(x)
    rr, cc = np.indices(x.shape)
    return np.where((rr + cc) & 1, 0, x).astype(x.dtype)


def black(x):
    """BlackArray2D.cpp:20-24: x where (column + row) is odd, else 0."""
    x = np.asarray(x)
    rr, cc = np.indices(x.shape)
    return np.where((rr + cc) & 1, x, 0).astype(x.dtype)
