"""Pins of the oracle's WarpImg2D (WarpImg2D.cpp:33-75).

The fp64 form is checked against scipy.ndimage.map_coordinates (order 1) where
the kernel's four taps are inside the image and the offsets are not integers
(there the kernel's split equals floor), against the exact reproduction of
affine images by bilinear interpolation, against whole-pixel shifts, and against
the kernel's own rule for negative whole offsets (truncate, then step down:
offset -1 samples x-2 with weight 0 and x-1 with weight 1, inside only if x-2 is);
the fp32 form against the fp64 one.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

import oracle


def inside(H, W, ox, oy):
    mx = np.trunc(ox).astype(np.int64)
    my = np.trunc(oy).astype(np.int64)
    mx -= (ox < 0)
    my -= (oy < 0)
    rr, cc = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    sx, sy = cc + mx, rr + my
    return (sx >= 0) & (sx + 1 < W) & (sy >= 0) & (sy + 1 < H)


@pytest.mark.parametrize("shape", [(9, 13), (32, 20)])
def test_matches_map_coordinates_inside(shape):
    rng = np.random.default_rng(5)
    H, W = shape
    src = rng.normal(size=(H, W))
    ox = rng.uniform(-4.5, 4.5, (H, W))
    oy = rng.uniform(-4.5, 4.5, (H, W))
    got = oracle.warp_img2d(src, ox, oy, np.float64)
    rr, cc = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    ref = ndi.map_coordinates(src, [rr + oy, cc + ox], order=1, mode="nearest")
    m = inside(H, W, ox, oy)
    assert m.sum() > H * W // 4
    assert np.allclose(got[m], ref[m], rtol=0, atol=1e-12)
    assert np.array_equal(got[~m], src[~m])   # outside: the source pixel itself


def test_affine_image_is_reproduced():
    H, W = 16, 24
    rr, cc = np.meshgrid(np.arange(H, dtype=float), np.arange(W, dtype=float), indexing="ij")
    f = lambda y, x: 0.3 + 1.7 * x - 0.6 * y  # noqa: E731
    rng = np.random.default_rng(6)
    ox = rng.uniform(-3, 3, (H, W))
    oy = rng.uniform(-3, 3, (H, W))
    got = oracle.warp_img2d(f(rr, cc), ox, oy, np.float64)
    m = inside(H, W, ox, oy)
    assert np.allclose(got[m], f(rr + oy, cc + ox)[m], rtol=0, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_whole_pixel_shifts(dtype):
    H, W = 7, 10
    src = np.arange(H * W, dtype=float).reshape(H, W)
    # +2 / +1: exact shift where the 2x2 neighbourhood fits
    got = oracle.warp_img2d(src, np.full((H, W), 2.0), np.full((H, W), 1.0), dtype)
    for r in range(H):
        for c in range(W):
            ok = c + 3 < W and r + 2 < H
            assert got[r, c] == (src[r + 1, c + 2] if ok else src[r, c])
    # -1 in x: the kernel steps down to x-2 with weight 0 and takes x-1
    got = oracle.warp_img2d(src, np.full((H, W), -1.0), np.zeros((H, W)), dtype)
    for r in range(H):
        for c in range(W):
            ok = c - 2 >= 0 and r + 1 < H
            assert got[r, c] == (src[r, c - 1] if ok else src[r, c])


def test_fp32_close_to_fp64():
    rng = np.random.default_rng(7)
    src = rng.uniform(0, 1, (40, 50)).astype(np.float32)
    ox = rng.uniform(-2.5, 2.5, (40, 50)).astype(np.float32)
    oy = rng.uniform(-2.5, 2.5, (40, 50)).astype(np.float32)
    a = oracle.warp_img2d(src, ox, oy)
    b = oracle.warp_img2d(src, ox, oy, np.float64)
    assert np.max(np.abs(a - b)) <= 4e-7
