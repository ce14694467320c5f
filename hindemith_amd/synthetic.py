"""Seeded synthetic inputs for the Horn-Schunck workloads.

Frames are sums of isotropic Gaussian blobs plus white noise; the second frame
is the first bilinearly warped by a known flow field, so the true flow is known
and an endpoint error can be reported.  Everything is numpy with
``default_rng(seed)`` so tests and benchmarks regenerate identical inputs.

The stencils are the Horn & Schunck (1981) choices a user passes to the
``@hm`` program as ``Stencil`` kwargs (offsets are (offx, offy) = (column, row),
as in Stencil2DMVM.cpp:40-41):

* ``Dx``: (0,0) -1/4, (1,0) +1/4, (0,1) -1/4, (1,1) +1/4
* ``Dy``: (0,0) -1/4, (0,1) +1/4, (1,0) -1/4, (1,1) +1/4
* ``Dt``: the same four taps, all +1/4 (the 2x2x2 cube; It = Dt*I2 - Dt*I1)
* ``Avg``: 4 edge neighbours 1/6, 4 corner neighbours 1/12
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "gaussian_blobs", "warp_bilinear", "constant_flow", "affine_flow",
    "hs_stencil_taps", "laplacian5_taps", "hs_pair", "batched_pairs", "laplacian5_dia_planes",
    "dia2d_to_csr", "powerlaw_csr",
]


def gaussian_blobs(H, W, rng, n_blobs=32, sigma=(8.0, 24.0), amp=(0.0, 1.0),
                   noise_std=0.01):
    """Sum of ``n_blobs`` isotropic Gaussians + N(0, noise_std) (std, not variance).

    Centres ~ U(image), sigma ~ U(sigma), amplitude ~ U(amp).  Separable
    evaluation (outer product per blob) keeps 4K frames to about a second.
    Returns float64.
    """
    img = np.zeros((H, W), dtype=np.float64)
    xs = np.arange(W, dtype=np.float64)
    ys = np.arange(H, dtype=np.float64)
    cx = rng.uniform(0, W, n_blobs)
    cy = rng.uniform(0, H, n_blobs)
    sg = rng.uniform(sigma[0], sigma[1], n_blobs)
    am = rng.uniform(amp[0], amp[1], n_blobs)
    for k in range(n_blobs):
        gx = np.exp(-((xs - cx[k]) ** 2) / (2.0 * sg[k] ** 2))
        gy = np.exp(-((ys - cy[k]) ** 2) / (2.0 * sg[k] ** 2))
        img += am[k] * np.outer(gy, gx)
    if noise_std > 0:
        img += rng.normal(0.0, noise_std, size=(H, W))
    return img


def warp_bilinear(img, fx, fy):
    """frame2(x) = frame1(x - f(x)), bilinear, replicate (clamped) border.

    ``fx`` (columns) and ``fy`` (rows) are scalars or H x W arrays.
    """
    H, W = img.shape
    yy, xx = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64),
                         indexing="ij")
    sx = np.clip(xx - fx, 0.0, W - 1.0)
    sy = np.clip(yy - fy, 0.0, H - 1.0)
    x0 = np.floor(sx).astype(np.int64)
    y0 = np.floor(sy).astype(np.int64)
    x1 = np.minimum(x0 + 1, W - 1)
    y1 = np.minimum(y0 + 1, H - 1)
    ax = sx - x0
    ay = sy - y0
    top = (1.0 - ax) * img[y0, x0] + ax * img[y0, x1]
    bot = (1.0 - ax) * img[y1, x0] + ax * img[y1, x1]
    return (1.0 - ay) * top + ay * bot


def constant_flow(H, W, fx, fy):
    return np.full((H, W), float(fx)), np.full((H, W), float(fy))


def affine_flow(H, W, rot_deg=0.5, scale=1.01, trans=(2.0, -1.0)):
    """f(x) = (scale*R(rot) - I)(x - c) + t about the image centre c."""
    th = math.radians(rot_deg)
    A = scale * np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]]) - np.eye(2)
    cx, cy = (W - 1) / 2.0, (H - 1) / 2.0
    yy, xx = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64),
                         indexing="ij")
    dx, dy = xx - cx, yy - cy
    fx = A[0, 0] * dx + A[0, 1] * dy + trans[0]
    fy = A[1, 0] * dx + A[1, 1] * dy + trans[1]
    return fx, fy


def hs_stencil_taps():
    """(weights f32, offx int32, offy int32) for Dx, Dy, Dt, Avg."""
    q = 0.25
    dx = ([-q, q, -q, q], [0, 1, 0, 1], [0, 0, 1, 1])
    dy = ([-q, q, -q, q], [0, 0, 1, 1], [0, 1, 0, 1])
    dt = ([q, q, q, q], [0, 1, 0, 1], [0, 0, 1, 1])
    e, c = 1.0 / 6.0, 1.0 / 12.0
    avg = ([e, e, e, e, c, c, c, c],
           [-1, 1, 0, 0, -1, 1, -1, 1],
           [0, 0, -1, 1, -1, -1, 1, 1])
    out = {}
    for name, (w, ox, oy) in (("Dx", dx), ("Dy", dy), ("Dt", dt), ("Avg", avg)):
        out[name] = (np.asarray(w, dtype=np.float32), np.asarray(ox, dtype=np.int32),
                     np.asarray(oy, dtype=np.int32))
    return out


def laplacian5_taps():
    """5-point Laplacian: centre 4, neighbours -1."""
    return (np.asarray([4, -1, -1, -1, -1], dtype=np.float32),
            np.asarray([0, -1, 1, 0, 0], dtype=np.int32),
            np.asarray([0, 0, 0, -1, 1], dtype=np.int32))


def hs_pair(H, W, seed=0, flow="constant", const_flow=(1.5, -0.75), trans=None):
    """Frame pair + true flow.  flow: 'constant' | 'affine' | 'translation'."""
    rng = np.random.default_rng(seed)
    f1 = gaussian_blobs(H, W, rng)
    if flow == "constant":
        fx, fy = constant_flow(H, W, *const_flow)
    elif flow == "affine":
        fx, fy = affine_flow(H, W)
    elif flow == "translation":
        tx, ty = trans if trans is not None else rng.uniform(-3.0, 3.0, 2)
        fx, fy = constant_flow(H, W, tx, ty)
    else:
        raise ValueError(flow)
    f2 = warp_bilinear(f1, fx, fy)
    return (f1.astype(np.float32), f2.astype(np.float32),
            fx.astype(np.float32), fy.astype(np.float32))


def batched_pairs(n_pairs=64, seed=0, side=(64, 1024), alpha=(0.5, 2.0)):
    """Uneven pairs: W, H ~ U(side) independently, translation ~ U(-3,3)^2,
    alpha ~ U(alpha).  Returns a list of dicts (I1, I2, fx, fy, alpha)."""
    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n_pairs):
        W = int(rng.integers(side[0], side[1] + 1))
        H = int(rng.integers(side[0], side[1] + 1))
        tx, ty = rng.uniform(-3.0, 3.0, 2)
        a = float(rng.uniform(*alpha))
        prng = np.random.default_rng([seed, i])
        f1 = gaussian_blobs(H, W, prng)
        f2 = warp_bilinear(f1, tx, ty)
        pairs.append(dict(I1=f1.astype(np.float32), I2=f2.astype(np.float32),
                          fx=np.float32(tx), fy=np.float32(ty), alpha=np.float32(a)))
    return pairs


def laplacian5_dia_planes(H, W):
    """The 5-point Laplacian as SpMat2D planes (constant per diagonal)."""
    w, ox, oy = laplacian5_taps()
    return np.stack([np.full((H, W), float(v), np.float32) for v in w]), ox, oy


def dia2d_to_csr(planes, offx, offy):
    """CSR (indptr, indices, data; int32/float32) of a 2D DIA matrix with the
    coefficient read at the neighbour index and out-of-range terms dropped.
    Entries of a row are ordered by diagonal, as the DIA kernel visits them."""
    nd, H, W = planes.shape
    n = H * W
    rr, cc = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    cols = np.full((n, nd), -1, np.int64)
    vals = np.zeros((n, nd), np.float32)
    for d in range(nd):
        r2, c2 = rr + int(offy[d]), cc + int(offx[d])
        m = ((r2 >= 0) & (r2 < H) & (c2 >= 0) & (c2 < W)).reshape(-1)
        q = (r2 * W + c2).reshape(-1)
        cols[m, d] = q[m]
        vals[m, d] = planes[d].reshape(-1)[q[m]]
    keep = cols >= 0
    indptr = np.zeros(n + 1, np.int64)
    indptr[1:] = np.cumsum(keep.sum(1))
    return (indptr.astype(np.int32), cols[keep].astype(np.int32), vals[keep].astype(np.float32))


def powerlaw_csr(n_rows, rng, s=1.5, lo=1, hi=512, n_cols=None):
    """Random CSR: row lengths Zipf(s) truncated to [lo, hi] (sampled from the
    normalised truncated distribution), columns U[0, n_cols), values N(0, 1)."""
    n_cols = n_rows if n_cols is None else n_cols
    k = np.arange(lo, hi + 1, dtype=np.float64)
    p = k ** -s
    p /= p.sum()
    lens = rng.choice(k.astype(np.int64), size=n_rows, p=p)
    indptr = np.zeros(n_rows + 1, np.int64)
    indptr[1:] = np.cumsum(lens)
    nnz = int(indptr[-1])
    indices = rng.integers(0, n_cols, nnz, dtype=np.int32)
    data = rng.standard_normal(nnz, dtype=np.float32)
    return indptr.astype(np.int32), indices, data
