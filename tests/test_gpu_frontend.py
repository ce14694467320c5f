"""The drop-in @hm path end to end on the GPU (record API of libhmx).

The canonical Horn-Schunck program must equal the fp32 oracle bit for bit
(both the fused HS loop and the generic program path perform the reference's
operations in the reference's order) and stay within 1e-4 of the fp64
textbook solution.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from hindemith_amd import synthetic as S

pytestmark = pytest.mark.gpu
TAPS = S.hs_stencil_taps()


@pytest.fixture(scope="module", autouse=True)
def env():
    os.environ["HM_GPU0"] = "1"
    from hindemith_amd import hindemith as H
    H.defines.update(nd20=1, nd21=1, nt20=8, nt21=8, ne20=2, ne21=2, nd1=1, nt1=128, ne1=2)
    return H


def kwargs(I1, I2, alpha2=1.0):
    from hindemith_amd.hindemith import Stencil
    z = np.zeros_like(I1)
    return dict(I1=I1, I2=I2, Dx=Stencil(*TAPS["Dx"]), Dy=Stencil(*TAPS["Dy"]),
                Dt=Stencil(*TAPS["Dt"]), Avg=Stencil(*TAPS["Avg"]),
                alpha2=np.float32(alpha2), u=z.copy(), v=z.copy())


def make_programs(H):
    hm, sqrt, sum2d = H.hm, H.sqrt, H.sum2d

    @hm
    def hs_canonical(I1, I2, Dx, Dy, Dt, Avg, alpha2, u, v):
        Ix = Dx * I1 + Dx * I2
        Iy = Dy * I1 + Dy * I2
        It = Dt * I2 - Dt * I1
        den = alpha2 + Ix * Ix + Iy * Iy
        for i in range(100):
            ubar = Avg * u
            vbar = Avg * v
            t = (Ix * ubar + Iy * vbar + It) / den
            u = ubar - Ix * t
            v = vbar - Iy * t
        return u, v

    @hm
    def hs_generic(I1, I2, Dx, Dy, Dt, Avg, alpha2, u, v):
        # the extra copy keeps the loop off the fused HS kernel: generic programs
        Ix = Dx * I1 + Dx * I2
        Iy = Dy * I1 + Dy * I2
        It = Dt * I2 - Dt * I1
        den = alpha2 + Ix * Ix + Iy * Iy
        for i in range(100):
            ubar = Avg * u
            vbar = Avg * v
            t = (Ix * ubar + Iy * vbar + It) / den
            u = ubar - Ix * t
            v = vbar - Iy * t
            last_t = t
        return u, v, last_t

    @hm
    def residual(I1, I2, Dx, Dy, Dt, Avg, alpha2, u, v):
        Ix = Dx * I1 + Dx * I2
        Iy = Dy * I1 + Dy * I2
        It = Dt * I2 - Dt * I1
        r = Ix * u + Iy * v + It
        res = sqrt(sum2d(r * r))
        return res

    return hs_canonical, hs_generic, residual


@pytest.fixture(scope="module")
def progs(env):
    return make_programs(env)


def test_canonical_hs_config1_bitexact(progs):
    hs_canonical = progs[0]
    I1, I2, fx, fy = S.hs_pair(256, 256, seed=0)
    kw = kwargs(I1, I2)
    u_in = kw["u"]
    u, v = hs_canonical(**kw)
    assert "hs_jacobi_fused x100" in hs_canonical.plan()
    ru, rv = oracle.hs_solve(I1, I2, 1.0, TAPS, 100)
    assert np.array_equal(u, ru) and np.array_equal(v, rv)
    assert u is u_in  # returned kwarg arrays are updated in place, as in the reference
    u64, v64 = oracle.np_hs_solve_f64(I1, I2, 1.0, TAPS, 100)
    assert np.max(np.abs(u - u64)) <= 1e-4 and np.max(np.abs(v - v64)) <= 1e-4


def test_canonical_hs_odd_shape_and_rerun(progs):
    hs_canonical = progs[0]
    I1, I2, _, _ = S.hs_pair(97, 203, seed=3)
    kw = kwargs(I1, I2, 0.5)
    u, v = hs_canonical(**kw)
    ru, rv = oracle.hs_solve(I1, I2, 0.5, TAPS, 100)
    assert np.array_equal(u, ru) and np.array_equal(v, rv)
    # second call with the (in-place updated) u, v continues from them
    u2, v2 = hs_canonical(**kw)
    ru2, rv2 = oracle.hs_solve(I1, I2, 0.5, TAPS, 100, ru, rv)
    assert np.array_equal(u2, ru2) and np.array_equal(v2, rv2)


def test_generic_program_path_bitexact(progs):
    hs_generic = progs[1]
    I1, I2, _, _ = S.hs_pair(64, 80, seed=4)
    u, v, t = hs_generic(**kwargs(I1, I2))
    assert "hs_jacobi_fused" not in hs_generic.plan() and "loop x100" in hs_generic.plan()
    ru, rv = oracle.hs_solve(I1, I2, 1.0, TAPS, 100)
    assert np.array_equal(u, ru) and np.array_equal(v, rv)
    assert t.shape == u.shape


def test_residual_scalar(progs):
    residual = progs[2]
    I1, I2, _, _ = S.hs_pair(50, 70, seed=6)
    kw = kwargs(I1, I2)
    rng = np.random.default_rng(0)
    kw["u"] = rng.normal(size=I1.shape).astype(np.float32)
    kw["v"] = rng.normal(size=I1.shape).astype(np.float32)
    res = residual(**kw)
    Ix, Iy, It = oracle.hs_setup(I1, I2, TAPS["Dx"], TAPS["Dy"], TAPS["Dt"])
    r = ((Ix * kw["u"]) + (Iy * kw["v"])) + It
    ref = np.sqrt(np.sum((r * r).astype(np.float64)))
    assert isinstance(res, float) and res == pytest.approx(ref, rel=1e-6)


def test_dia2d_and_csr_programs(env):
    hm = env.hm

    @hm
    def dia(M, x):
        y = M * x
        return y

    @hm
    def csr(A, x):
        y = A * x
        return y

    rng = np.random.default_rng(1)
    w, ox, oy = S.laplacian5_taps()
    planes = rng.normal(size=(5, 40, 33)).astype(np.float32)
    x = rng.uniform(-1, 1, (40, 33)).astype(np.float32)
    y = dia(M=env.SpMat2D(planes, ox, oy), x=x)
    assert np.array_equal(y, oracle.dia2d(planes, ox, oy, x))
    A = sp.random(500, 400, density=0.03, random_state=2, format="csr", dtype=np.float32)
    xv = rng.uniform(-1, 1, 400).astype(np.float32)
    yv = csr(A=A, x=xv)
    ref = A.astype(np.float64) @ xv.astype(np.float64)
    assert yv.shape == (500,) and np.allclose(yv, ref, rtol=1e-5, atol=1e-5)


def test_scalar_ops_and_column_major(env):
    hm, sqrt = env.hm, env.sqrt

    @hm
    def scal(a, b, x):
        s = a * b
        t = sqrt(s)
        y = x * t
        return y, t

    x = np.asfortranarray(np.random.default_rng(2).normal(size=(7, 5)).astype(np.float32))
    y, t = scal(a=np.float32(2.0), b=np.float32(8.0), x=x)
    assert t == 4.0
    assert y.flags["F_CONTIGUOUS"] and np.array_equal(y, x * np.float32(4.0))


def test_stencil_program_bitexact_wide(env):
    hm = env.hm

    @hm
    def smooth(S5, x):
        for i in range(3):
            x = S5 * x
        return x

    st = (np.array([0.2, 0.2, 0.2, 0.2, 0.2], np.float32), np.array([0, -2, 2, 0, 0], np.int32),
          np.array([0, 0, 0, -2, 2], np.int32))
    x0 = np.random.default_rng(3).normal(size=(30, 41)).astype(np.float32)
    y = smooth(S5=env.Stencil(*st), x=x0.copy())
    ref = x0
    for _ in range(3):
        ref = oracle.stencil2d(ref, st)
    assert np.array_equal(y, ref)


def test_dia_matrix_residual_program():
    """scipy dia_matrix * Array through the @hm path (1-D DIA fused into the
    elementwise program) against the oracle."""
    from hindemith_amd.hindemith import defines, hm
    defines.update(nd1=1, nt1=128, ne1=2)

    @hm
    def residual(A, x, b):
        r = b - A * x
        return r
    rng = np.random.default_rng(1)
    n = 5000
    offsets = np.array([0, -1, 1, -70, 70], np.int32)
    data = rng.normal(size=(5, n)).astype(np.float32)
    A = sp.dia_matrix((data, offsets), shape=(n, n))
    x = rng.normal(size=n).astype(np.float32)
    b = rng.normal(size=n).astype(np.float32)
    r = residual(A=A, x=x, b=b)
    ref = (b - oracle.dia1d(data, offsets, x)).astype(np.float32)
    assert np.array_equal(r, ref)
