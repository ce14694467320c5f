"""Generated program kernels against the interpreter, on the GPU.

Every fused group of a record runs as a kernel generated for it and compiled
with hipRTC; with HMX_PROGRAM_INTERP=1 set when the record is compiled, the
same groups run on the interpreter (program_kernel).  Both perform the same
fp32 operation per instruction in the same order, the same clamped fma chain
per stencil, the same skipped-tap DIA rule and the same reduction tree, so
their results must be identical bit for bit -- including reductions, whose
partials come from the same grid -- on programs that use every instruction
kind.  The oracle comparisons of the other @hm tests run on the generated
kernels too.
"""
import os

import numpy as np
import pytest

from hindemith_amd import synthetic as S

pytestmark = pytest.mark.gpu
TAPS = S.hs_stencil_taps()


@pytest.fixture(scope="module")
def H():
    os.environ["HM_GPU0"] = "1"
    from hindemith_amd import hindemith as Hm
    Hm.defines.update(nd20=1, nd21=1, nt20=8, nt21=8, ne20=2, ne21=2, nd1=1, nt1=128, ne1=2)
    return Hm


def both_ways(make, kwargs):
    """Run the program made by make() compiled on the interpreter, then again
    (a new function object, so a new record) on generated kernels."""
    os.environ["HMX_PROGRAM_INTERP"] = "1"
    try:
        f_int = make()
        out_int = f_int(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                           for k, v in kwargs.items()})
    finally:
        del os.environ["HMX_PROGRAM_INTERP"]
    f_jit = make()
    out_jit = f_jit(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                       for k, v in kwargs.items()})
    assert "extern \"C\" __global__" in f_jit.program_source()
    return out_int, out_jit


def same_bits(a, b):
    a, b = np.atleast_1d(np.asarray(a)), np.atleast_1d(np.asarray(b))
    assert a.shape == b.shape and a.dtype == b.dtype
    assert np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                          np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize("shape", [(1, 1), (7, 300), (129, 257), (1080, 1920), (4100, 64)])
def test_stencils_division_sqrt_and_reduction(H, shape):
    def make():
        @H.hm
        def prog(I1, I2, Dx, Dy, Dt, alpha2, u, v):
            Ix = Dx * I1 + Dx * I2
            Iy = Dy * I1 + Dy * I2
            It = Dt * I2 - Dt * I1
            den = alpha2 + Ix * Ix + Iy * Iy
            r = (Ix * u + Iy * v + It) / den
            q = sqrt(den) - r  # noqa: F821
            res = sqrt(sum2d(r * r))  # noqa: F821
            return Ix, Iy, It, den, q, res
        return prog
    rng = np.random.default_rng(sum(shape))
    I1 = rng.standard_normal(shape).astype(np.float32)
    I2 = rng.standard_normal(shape).astype(np.float32)
    kw = dict(I1=I1, I2=I2, Dx=H.Stencil(*TAPS["Dx"]), Dy=H.Stencil(*TAPS["Dy"]),
              Dt=H.Stencil(*TAPS["Dt"]), alpha2=np.float64(0.25),
              u=rng.standard_normal(shape).astype(np.float32),
              v=rng.standard_normal(shape).astype(np.float32))
    a, b = both_ways(make, kw)
    for x, y in zip(a, b):
        same_bits(x, y)


def test_dia_red_black_and_one_dimensional(H):
    def make_dia():
        @H.hm
        def prog(M, x, y):
            z = M * x
            w = red(z) + black(y)  # noqa: F821
            return z, w
        return prog
    w, ox, oy = S.laplacian5_taps()
    rng = np.random.default_rng(5)
    planes = rng.standard_normal((5, 33, 65)).astype(np.float32)
    kw = dict(M=H.SpMat2D(planes, ox, oy), x=rng.standard_normal((33, 65)).astype(np.float32),
              y=rng.standard_normal((33, 65)).astype(np.float32))
    a, b = both_ways(make_dia, kw)
    for x, y in zip(a, b):
        same_bits(x, y)

    def make_1d():
        @H.hm
        def prog(a, x, y, n):
            z = a * x + y
            w = sqrt(z * z) - n  # noqa: F821
            s = sum(w)  # noqa: F821
            return w, s
        return prog
    x = rng.standard_normal(100_003).astype(np.float32)
    kw = dict(a=np.float32(1.5), x=x, y=rng.standard_normal(100_003).astype(np.float32),
              n=np.int64(3))
    a, b = both_ways(make_1d, kw)
    for x, y in zip(a, b):
        same_bits(x, y)
