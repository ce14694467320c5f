"""Parity at the benchmark's own sizes: the Horn-Schunck solve at 1920x1080 and
3840x2160 for 500 iterations, on the affine-flow inputs the benchmark times.

The fused kernel is compared bit for bit with the fp32 oracle (the canonical
program's operations in the reference's order, Templates.py:57-71 with the
taps of Stencil2DMVM.cpp:36-47), through every path that reaches it at this
size: ``hmx_hs_solve`` with the default configuration (variant 24, k = 8, the
one the benchmark's tuner keeps), explicit variants of the other tile-body
forms through ``hmx_hs_jacobi_cfg``, and the drop-in ``@hm`` canonical
program.  The north-star tolerance, max |du|, |dv| <= 1e-4 against the fp64
oracle, is checked on the same results, and the endpoint error against the
true synthetic flow is printed.
"""
import os

import numpy as np
import pytest

import oracle
from hindemith_amd import synthetic as S

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TAPS = S.hs_stencil_taps()
ITERS = 500
TOL = 1e-4  # max |du|, |dv| against the fp64 oracle after 500 iterations


class _Case:
    """Inputs of one benchmark size and the oracle's fp32 and fp64 solutions,
    computed once per module (the fp32 solve at 4K takes seconds on the host)."""

    def __init__(self, H, W, seed):
        self.H, self.W = H, W
        self.I1, self.I2, self.fx, self.fy = S.hs_pair(H, W, seed=seed, flow="affine")
        self._f32 = self._f64 = None

    def f32(self):
        if self._f32 is None:
            self._f32 = oracle.hs_solve(self.I1, self.I2, 1.0, TAPS, ITERS)
        return self._f32

    def f64(self):
        if self._f64 is None:
            Ix, Iy, It = oracle.hs_setup(self.I1, self.I2, TAPS["Dx"], TAPS["Dy"], TAPS["Dt"],
                                         dtype=np.float64)
            self._f64 = oracle.hs_jacobi(Ix, Iy, It, 1.0, TAPS["Avg"], ITERS, dtype=np.float64)
        return self._f64


# the benchmark's inputs: seed 0 at 1080p, seed 1 at 4K (bench.py)
_CASES = {}


def case(H, W):
    if (H, W) not in _CASES:
        _CASES[(H, W)] = _Case(H, W, 0 if (H, W) == (1080, 1920) else 1)
    return _CASES[(H, W)]


@pytest.fixture(scope="module")
def K():
    from hindemith_amd import kernels
    return kernels


def _check(c, u, v, what):
    ru, rv = c.f32()
    assert np.array_equal(u, ru), f"{what}: u differs from the fp32 oracle"
    assert np.array_equal(v, rv), f"{what}: v differs from the fp32 oracle"
    u64, v64 = c.f64()
    du, dv = float(np.max(np.abs(u - u64))), float(np.max(np.abs(v - v64)))
    assert du <= TOL and dv <= TOL, (what, du, dv)
    epe = oracle.endpoint_error(u, v, c.fx, c.fy)
    print(f"{what}: bit-exact vs fp32 oracle; max |du| {du:.3g}, |dv| {dv:.3g} vs fp64; "
          f"EPE {epe:.4f} px")


@pytest.mark.parametrize("shape", [(1080, 1920), (2160, 3840)])
def test_hs_solve_default_config(K, shape):
    """hmx_hs_solve (setup + 500 fused iterations) with the default configuration."""
    c = case(*shape)
    I1, I2 = torch.from_numpy(c.I1).cuda(), torch.from_numpy(c.I2).cuda()
    u, v = K.hs_solve(I1, I2, 1.0, TAPS, ITERS)
    torch.cuda.synchronize()
    _check(c, u.cpu().numpy(), v.cpu().numpy(), f"hs_solve {shape[1]}x{shape[0]}")


# One variant per tile-body form: 24 the default (double-buffered, reciprocal
# quotient), 27 its ring-ordered form, 28 the persistent launch of 24's tile
# body, 17 double-buffered with the IEEE division, 3 the single-window
# two-barrier step, 12 the column-strip body, 0 one plain sweep per launch.
@pytest.mark.parametrize("variant,k", [(24, 8), (24, 7), (27, 8), (27, 5), (28, 8), (28, 5),
                                       (17, 8), (3, 8), (12, 8), (0, 1)])
def test_hs_jacobi_1080p_variants(K, variant, k):
    c = case(1080, 1920)
    I1, I2 = torch.from_numpy(c.I1).cuda(), torch.from_numpy(c.I2).cuda()
    Ix, Iy, It = K.hs_setup(I1, I2, TAPS["Dx"], TAPS["Dy"], TAPS["Dt"])
    u = torch.zeros_like(I1)
    v = torch.zeros_like(I1)
    K.hs_jacobi(Ix, Iy, It, 1.0, TAPS["Avg"], ITERS, u, v, variant=variant, k=k)
    torch.cuda.synchronize()
    _check(c, u.cpu().numpy(), v.cpu().numpy(), f"variant {variant} k {k} 1920x1080")


def test_hm_canonical_program_1080p():
    """The drop-in @hm canonical program at 1920x1080, 500 iterations, host arrays."""
    os.environ["HM_GPU0"] = "1"
    from hindemith_amd import hindemith as Hm
    Hm.defines.update(nd20=1, nd21=1, nt20=8, nt21=8, ne20=2, ne21=2)

    @Hm.hm
    def hs_canonical_500(I1, I2, Dx, Dy, Dt, Avg, alpha2, u, v):
        Ix = Dx * I1 + Dx * I2
        Iy = Dy * I1 + Dy * I2
        It = Dt * I2 - Dt * I1
        den = alpha2 + Ix * Ix + Iy * Iy
        for i in range(500):
            ubar = Avg * u
            vbar = Avg * v
            t = (Ix * ubar + Iy * vbar + It) / den
            u = ubar - Ix * t
            v = vbar - Iy * t
        return u, v

    c = case(1080, 1920)
    z = np.zeros_like(c.I1)
    st = {n: Hm.Stencil(*TAPS[n]) for n in ("Dx", "Dy", "Dt", "Avg")}
    u, v = hs_canonical_500(I1=c.I1, I2=c.I2, alpha2=np.float32(1.0), u=z, v=z.copy(), **st)
    assert "hs_jacobi_fused x500" in hs_canonical_500.plan()
    _check(c, u, v, "@hm canonical 1920x1080")
