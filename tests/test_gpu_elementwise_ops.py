"""Elementwise and reduction operations through the drop-in @hm path on the GPU:
Array2DSqrt, Array2DScalarOp / ScalarArray2DOp, and the 1-D ArraySqrt,
ArrayScalarOp / ScalarArrayOp and ArraySum, alone and fused into one group.

Each elementwise operation is one IEEE fp32 op with a correctly rounded sqrt
(Array2DSqrt.cpp:26, ArraySqrt.cpp:26, Array2DScalarOp.cpp:32,
ScalarArray2DOp.cpp:32, ArrayScalarOp.cpp:32, ScalarArrayOp.cpp:32), so the
results equal numpy float32 evaluated one operator at a time, bit for bit.
ArraySum (ArraySum.cpp:27-38) accumulates in fp64 here: its value is within
one fp32 rounding of the exact sum.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

F = np.float32


@pytest.fixture(scope="module")
def Hm():
    os.environ["HM_GPU0"] = "1"
    from hindemith_amd import hindemith as H
    H.defines.update(nd20=1, nd21=1, nt20=8, nt21=8, ne20=2, ne21=2, nd1=1, nt1=128, ne1=2)
    return H


@pytest.fixture(scope="module")
def progs(Hm):
    hm, sqrt, sum2d = Hm.hm, Hm.sqrt, Hm.sum2d

    @hm
    def sqrt2d(x):
        y = sqrt(x)
        return y

    @hm
    def fused2d(a, b, s):
        y = sqrt(a * a + b) * s
        z = s - sqrt(b)
        w = a / s + s * b
        return y, z, w

    @hm
    def norm2d(a, b):
        r = a - b
        n = sqrt(sum2d(r * r))
        return n

    @hm
    def sqrt1d(x):
        y = sqrt(x)
        return y

    @hm
    def scal1d(x, s):
        y = x * s
        z = s * x
        w = s - x
        d = x / s
        return y, z, w, d

    @hm
    def sum1d(x):
        t = sum(x * x)
        return t

    @hm
    def fused1d(x, b, s):
        r = b - x * s
        q = sqrt(r * r + b * b)
        n = sqrt(sum(r * r))
        return q, n

    return dict(sqrt2d=sqrt2d, fused2d=fused2d, norm2d=norm2d, sqrt1d=sqrt1d, scal1d=scal1d,
                sum1d=sum1d, fused1d=fused1d)


def _same(a, b):
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def _sum_close(got, vals):
    """fp64 accumulation then one rounding to fp32: within 2^-23 relative of the
    exact sum (plus a tiny absolute term for sums near zero)."""
    ref = float(np.sum(np.asarray(vals, np.float64)))
    return abs(got - ref) <= abs(ref) * 2.0 ** -23 + 1e-30 * len(vals)


@pytest.mark.parametrize("shape", [(1, 1), (1, 77), (63, 1), (37, 41), (1080, 1920)])
def test_array2d_sqrt(progs, shape):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    x = rng.uniform(0, 100, shape).astype(F)
    x.flat[0] = 0.0                      # exact zero
    if x.size > 2:
        x.flat[1] = F(2.0)               # irrational root
        x.flat[2] = F(1e-40)             # subnormal
    y = progs["sqrt2d"](x=x)
    assert _same(y, np.sqrt(x))


@pytest.mark.parametrize("shape", [(1, 1), (37, 41), (257, 130), (1080, 1920)])
def test_array2d_scalar_ops_fused(progs, shape):
    rng = np.random.default_rng(3)
    a = rng.normal(size=shape).astype(F)
    b = rng.uniform(0, 4, shape).astype(F)
    s = F(1.7)
    y, z, w = progs["fused2d"](a=a, b=b, s=s)
    assert _same(y, np.sqrt(a * a + b) * s)        # Array2DOp, Array2DSqrt, Array2DScalarOp
    assert _same(z, s - np.sqrt(b))                # ScalarArray2DOp
    assert _same(w, a / s + s * b)


def test_array2d_sum_then_scalar_sqrt(progs):
    rng = np.random.default_rng(4)
    a = rng.normal(size=(333, 517)).astype(F)
    b = rng.normal(size=(333, 517)).astype(F)
    n = progs["norm2d"](a=a, b=b)
    r = a - b
    assert isinstance(n, float)
    ref = np.sqrt(np.sum((r * r).astype(np.float64)))
    assert n == pytest.approx(ref, rel=2e-7)


@pytest.mark.parametrize("n", [1, 2, 1000, 4097, (1 << 20) + 3])
def test_array_sqrt(progs, n):
    x = np.random.default_rng(n).uniform(0, 50, n).astype(F)
    x[0] = 0.0
    y = progs["sqrt1d"](x=x)
    assert _same(y, np.sqrt(x))


@pytest.mark.parametrize("n", [1, 1000, 4097, (1 << 20) + 3])
def test_array_scalar_ops(progs, n):
    x = np.random.default_rng(n + 1).normal(size=n).astype(F)
    s = F(-0.37)
    y, z, w, d = progs["scal1d"](x=x, s=s)
    assert _same(y, x * s)   # ArrayScalarOp
    assert _same(z, s * x)   # ScalarArrayOp
    assert _same(w, s - x)
    assert _same(d, x / s)


@pytest.mark.parametrize("n", [1, 1000, 4097, (1 << 22) + 5])
def test_array_sum(progs, n):
    x = np.random.default_rng(n + 2).normal(size=n).astype(F)
    t = progs["sum1d"](x=x)
    assert isinstance(t, float)
    assert _sum_close(t, x * x), (t, float(np.sum((x * x).astype(np.float64))))


def test_array_sum_all_ones_exact(progs):
    n = (1 << 20) + 7
    t = progs["sum1d"](x=np.ones(n, F))
    assert t == float(n)


@pytest.mark.parametrize("n", [3, 5000, (1 << 20) + 1])
def test_array_fused_group(progs, n):
    rng = np.random.default_rng(n + 3)
    x = rng.normal(size=n).astype(F)
    b = rng.normal(size=n).astype(F)
    s = F(0.5)
    q, nrm = progs["fused1d"](x=x, b=b, s=s)
    r = b - x * s
    assert _same(q, np.sqrt(r * r + b * b))
    ref = np.sqrt(np.sum((r * r).astype(np.float64)))
    assert nrm == pytest.approx(ref, rel=2e-7)
