"""WarpImg2D on the GPU, bit-exact against the fp32 oracle (the generated
code's statement types: double weights, float products, fma accumulation),
and the warp inside an @hm program."""
import os

import numpy as np
import pytest

import oracle
from hindemith_amd import synthetic as S

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def offsets(rng, H, W, scale):
    ox = rng.uniform(-scale, scale, (H, W)).astype(np.float32)
    oy = rng.uniform(-scale, scale, (H, W)).astype(np.float32)
    # whole offsets of both signs and zero, and a few far outside the image
    ox.flat[::11] = np.round(ox.flat[::11])
    oy.flat[::13] = np.round(oy.flat[::13])
    ox.flat[::29] = 3.0 * max(H, W)
    oy.flat[::31] = -3.0 * max(H, W)
    return ox, oy


@pytest.mark.parametrize("shape", [(1, 1), (1, 40), (33, 1), (2, 2), (45, 77), (256, 300)])
@pytest.mark.parametrize("scale", [0.75, 6.0])
def test_warp_bitexact(shape, scale):
    from hindemith_amd import kernels as K
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    H, W = shape
    src = rng.normal(size=(H, W)).astype(np.float32)
    ox, oy = offsets(rng, H, W, scale)
    got = K.warp_img2d(dev(src), dev(ox), dev(oy)).cpu().numpy()
    assert np.array_equal(got, oracle.warp_img2d(src, ox, oy))


def test_warp_recovers_a_translated_frame():
    """Warping frame 2 of a constant-flow pair back by the true flow reproduces
    frame 1 in the interior to the synthetic noise level."""
    from hindemith_amd import kernels as K
    I1, I2, fx, fy = S.hs_pair(120, 160, seed=2, flow="constant")
    w = K.warp_img2d(dev(I2), dev(fx), dev(fy)).cpu().numpy()
    sl = (slice(8, -8), slice(8, -8))
    assert np.sqrt(np.mean((w[sl] - I1[sl]) ** 2)) < 0.05


def test_hm_warp_program_bitexact():
    os.environ["HM_GPU0"] = "1"
    from hindemith_amd.hindemith import defines, hm
    saved = dict(defines)
    defines.update(nd20=1, nd21=1, nt20=8, nt21=8, ne20=2, ne21=2)
    try:
        @hm
        def warp_residual(I1, I2, u, v):
            I2w = warp_img2d(I2, u, v)  # noqa: F821
            r = I2w - I1
            return r, I2w

        rng = np.random.default_rng(4)
        I1, I2 = (rng.normal(size=(37, 53)).astype(np.float32) for _ in range(2))
        u, v = offsets(rng, 37, 53, 2.5)
        r, I2w = warp_residual(I1=I1, I2=I2, u=u, v=v)
        ref = oracle.warp_img2d(I2, u, v)
        assert np.array_equal(I2w, ref)
        assert np.array_equal(r, (ref - I1).astype(np.float32))
    finally:
        defines.clear()
        defines.update(saved)
