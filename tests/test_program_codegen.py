"""Generated kernels of fused programs, on the CPU: each program's HIP source is
generated from its instruction list and compiled with hipRTC for gfx950 (no GPU
needed), so every instruction kind is known to build before it runs.  The GPU
tests run the same kernels and compare them bit for bit with the oracle and
with the interpreter."""
import numpy as np
import pytest

from hindemith_amd import frontend as fe
from hindemith_amd import synthetic as S
from hindemith_amd.hindemith import SpMat2D, Stencil, defines, hm, sum2d

TAPS = S.hs_stencil_taps()


@pytest.fixture(autouse=True)
def plan_only():
    fe.set_plan_only(True)
    saved = dict(defines)
    defines.clear()
    defines.update(nd20=1, nd21=1, nt20=8, nt21=8, ne20=2, ne21=2, nd1=1, nt1=128, ne1=2)
    yield
    defines.clear()
    defines.update(saved)
    fe.set_plan_only(False)


@hm
def hs_setup_and_residual(I1, I2, Dx, Dy, Dt, alpha2, u, v):
    Ix = Dx * I1 + Dx * I2
    Iy = Dy * I1 + Dy * I2
    It = Dt * I2 - Dt * I1
    den = alpha2 + Ix * Ix + Iy * Iy
    r = (Ix * u + Iy * v + It) / den
    res = sqrt(sum2d(r * r))
    return den, res


def test_stencils_division_reduction_and_scalars_build():
    z = np.zeros((24, 40), np.float32)
    hs_setup_and_residual.compile(I1=z, I2=z.copy(), Dx=Stencil(*TAPS["Dx"]),
                                  Dy=Stencil(*TAPS["Dy"]), Dt=Stencil(*TAPS["Dt"]),
                                  alpha2=np.float64(0.5), u=z.copy(), v=z.copy())
    src = hs_setup_and_residual.program_source(compile=True)
    assert src.count("extern \"C\" __global__") >= 1
    assert ("s = fmaf(w0_0, *(const float *)((const char *)(x + (long long)min(max(y0 + oy0_0, 0)"
            in src)
    assert "(float)*(const double *)a.scal[0]" in src       # a float64 scalar kwarg
    assert "acc0 += (double)" in src and "__shfl_xor" in src  # the sum2d partials
    assert " / " in src                                      # correctly rounded division
    assert "reg[" not in src                                 # no LDS register file


def test_dia_operand_and_sqrt_build():
    @hm
    def dia_sqrt(M, x):
        y = M * x
        z = sqrt(y * y + x)
        return z
    w, ox, oy = S.laplacian5_taps()
    dia_sqrt.compile(M=SpMat2D(np.ones((5, 8, 9), np.float32), ox, oy),
                     x=np.zeros((8, 9), np.float32))
    src = dia_sqrt.program_source(compile=True)
    assert "oy0_4 = a.opnd[0].oy[4]" in src and "oy0_5" not in src  # 5 diagonals
    assert "a.opnd[0].w[4 * ((long long)a.H * a.W) + q]" in src and "sqrtf(" in src


def test_one_dimensional_arrays_and_integer_scalars_build():
    @hm
    def saxpy(a, x, y, n):
        z = a * x + y
        w = z - n
        s = sum(w)
        return w, s
    x = np.zeros(1000, np.float32)
    saxpy.compile(a=np.float32(2.0), x=x, y=x.copy(), n=np.int64(3))
    src = saxpy.program_source(compile=True)
    assert "(float)*(const long long *)a.scal[" in src


def test_rows_per_thread():
    """A program without a reduction handles four rows per thread, its stores
    after every row's loads; one with a reduction keeps one row per thread."""
    @hm
    def axpy(a, x, y):
        z = a * x + y
        return z
    x = np.zeros((10, 300), np.float32)
    axpy.compile(a=np.float32(2.0), x=x, y=x.copy())
    src = axpy.program_source(compile=True)
    assert "const int y3 = min(ybase + 3, a.H - 1);" in src and "ybase = blockIdx.y * 4" in src
    assert src.index("st0_3 = ") < src.index("__hip_atomic_store((float *)((char *)(a.out[0] + (long long)y0")
    z = np.zeros((24, 40), np.float32)
    hs_setup_and_residual.compile(I1=z, I2=z.copy(), Dx=Stencil(*TAPS["Dx"]),
                                  Dy=Stencil(*TAPS["Dy"]), Dt=Stencil(*TAPS["Dt"]),
                                  alpha2=np.float32(0.5), u=z.copy(), v=z.copy())
    src = hs_setup_and_residual.program_source()
    assert "ybase = blockIdx.y * 1;" in src and "const int y1 =" not in src


def test_generated_source_is_deterministic():
    z = np.zeros((16, 16), np.float32)
    kw = dict(I1=z, I2=z.copy(), Dx=Stencil(*TAPS["Dx"]), Dy=Stencil(*TAPS["Dy"]),
              Dt=Stencil(*TAPS["Dt"]), alpha2=np.float32(1.0), u=z.copy(), v=z.copy())
    hs_setup_and_residual.compile(**kw)
    a = hs_setup_and_residual.program_source()
    hs_setup_and_residual.compile(**kw)
    assert hs_setup_and_residual.program_source() == a
