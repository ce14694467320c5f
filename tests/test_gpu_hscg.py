"""Preconditioned-CG Horn-Schunck operations on the GPU, bit-exact against the
fp32 oracle (same per-op rounding, fmas for the "+=" statements of
HSMatrixPowers), on odd shapes, degenerate 1-pixel dimensions, the singular
pass-through branch, and power counts that span several launches."""
import numpy as np
import pytest

import oracle
from hindemith_amd import synthetic as S

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 37), (29, 1), (33, 47), (70, 200), (257, 129)]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.fixture(scope="module")
def K():
    from hindemith_amd import kernels
    return kernels


@pytest.mark.parametrize("shape", SHAPES)
def test_precondition_bitexact(K, shape):
    rng = np.random.default_rng(hash(shape) % 1000)
    H, W = shape
    diag0 = rng.uniform(0.5, 2.0, (H, W)).astype(np.float32)
    diag1 = rng.uniform(0.5, 2.0, (H, W)).astype(np.float32)
    off = rng.uniform(-1.0, 1.0, (H, W)).astype(np.float32)
    x0, x1 = rng.normal(size=(2, H, W)).astype(np.float32)
    # singular blocks (a = d = b = 2 exactly) on every 7th pixel: pass-through branch
    lap = S.laplacian5_taps()
    s2 = float(np.float32(lap[0][2]))
    n = np.zeros((H, W), np.float32)
    n[:, :-1] += 1
    n[:, 1:] += 1
    n[:-1, :] += 1
    n[1:, :] += 1
    sing = (np.arange(H * W).reshape(H, W) % 7) == 0
    if abs(s2) in (0.25, 0.5, 1.0):
        diag0[sing] = 2.0 - n[sing] * s2
        diag1[sing] = diag0[sing]
        off[sing] = 2.0
    y0, y1 = K.hs_precondition(lap, dev(diag0), dev(diag1), dev(off), dev(x0), dev(x1))
    r0, r1 = oracle.hs_precondition(s2, diag0, diag1, off, x0, x1)
    assert np.array_equal(y0.cpu().numpy(), r0) and np.array_equal(y1.cpu().numpy(), r1)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("steps", [1, 2, 5, 9, 12, 20])
def test_matrix_powers_bitexact(K, shape, steps):
    rng = np.random.default_rng(steps * 31 + shape[0])
    H, W = shape
    Ix, Iy, r0, r1 = (rng.normal(size=(H, W)).astype(np.float32) for _ in range(4))
    lam = 0.35
    V0, V1 = K.hs_matrix_powers(lam, dev(Ix), dev(Iy), dev(r0), dev(r1), steps)
    R0, R1 = oracle.hs_matrix_powers(lam, Ix, Iy, r0, r1, steps)
    g0, g1 = V0.cpu().numpy(), V1.cpu().numpy()
    for s in range(steps):
        assert np.array_equal(g0[s], R0[s]) and np.array_equal(g1[s], R1[s]), s


def test_matrix_powers_on_hs_derivatives(K):
    """The operator of a real HS problem (1080p-like aspect, derivative cube of a
    blob pair) for 8 powers: bit-exact, and the powers stay finite."""
    taps = S.hs_stencil_taps()
    I1, I2, _, _ = S.hs_pair(180, 320, seed=3, flow="affine")
    Ix, Iy, It = oracle.hs_setup(I1, I2, taps["Dx"], taps["Dy"], taps["Dt"])
    r0 = -(Ix * It).astype(np.float32)
    r1 = -(Iy * It).astype(np.float32)
    V0, V1 = K.hs_matrix_powers(1.0, dev(Ix), dev(Iy), dev(r0), dev(r1), 8)
    R0, R1 = oracle.hs_matrix_powers(1.0, Ix, Iy, r0, r1, 8)
    assert np.array_equal(V0.cpu().numpy(), R0) and np.array_equal(V1.cpu().numpy(), R1)
    assert np.isfinite(R0).all()


def test_bad_arguments_raise(K):
    from hindemith_amd._lib import HindemithError
    x = dev(np.zeros((4, 5)))
    with pytest.raises(HindemithError):
        K.hs_precondition(([1.0, 2.0], [0, 1], [0, 0]), x, x, x, x, x)  # < 3 taps
    with pytest.raises(HindemithError):
        K.hs_matrix_powers(1.0, x, x, x, x, 0)


def test_hm_program_bitexact():
    """applyHSPreconditioner and hs_matrix_powers through the drop-in @hm path:
    the preconditioned residual and its STEPS powers equal the oracle chain."""
    import os
    os.environ["HM_GPU0"] = "1"
    from hindemith_amd.hindemith import (Stencil, applyHSPreconditioner, defines, hm,
                                         hs_matrix_powers)
    saved = dict(defines)
    defines.update(nd20=1, nd21=1, nt20=8, nt21=8, ne20=2, ne21=2, STEPS=6)
    try:
        @hm
        def cg_step(D, Ix, Iy, diag0, diag1, offdiag, r0, r1, lam):
            z0, z1 = applyHSPreconditioner(D, diag0, diag1, offdiag, r0, r1)
            V0, V1 = hs_matrix_powers(lam, Ix, Iy, z0, z1)
            return V0, V1, z0

        rng = np.random.default_rng(9)
        H, W = 45, 77
        Ix, Iy, r0, r1 = (rng.normal(size=(H, W)).astype(np.float32) for _ in range(4))
        diag0 = (Ix * Ix).astype(np.float32)
        diag1 = (Iy * Iy).astype(np.float32)
        off = (Ix * Iy).astype(np.float32)
        lap = S.laplacian5_taps()
        V0, V1, z0 = cg_step(D=Stencil(*lap), Ix=Ix, Iy=Iy, diag0=diag0, diag1=diag1,
                             offdiag=off, r0=r0, r1=r1, lam=np.float32(0.8))
        zr0, zr1 = oracle.hs_precondition(float(lap[0][2]), diag0, diag1, off, r0, r1)
        R0, R1 = oracle.hs_matrix_powers(0.8, Ix, Iy, zr0, zr1, 6)
        assert V0.shape == (6, H, W) and V1.shape == (6, H, W)
        assert np.array_equal(z0, zr0)
        assert np.array_equal(V0, R0) and np.array_equal(V1, R1)
        assert "hs_matrix_powers x6" in cg_step.plan()
    finally:
        defines.clear()
        defines.update(saved)
