"""Brox SOR building blocks on the GPU, bit-exact against the fp32 oracle, and
a red-black preconditioned sweep through the @hm path."""
import os

import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
OFFX = np.array([0, -1, 0, 1, 0], np.int32)
OFFY = np.array([-1, 0, 0, 0, 1], np.int32)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.fixture(scope="module")
def K():
    from hindemith_amd import kernels
    return kernels


@pytest.mark.parametrize("shape", [(1, 1), (1, 30), (25, 1), (41, 67), (300, 257)])
def test_brox_set_matrix_bitexact(K, shape):
    rng = np.random.default_rng(shape[0] + 3 * shape[1])
    H, W = shape
    psi = rng.uniform(0.1, 2.0, (H, W)).astype(np.float32)
    A0 = rng.normal(size=(5, H, W)).astype(np.float32)   # entries not written keep these
    A = dev(A0)
    K.brox_set_matrix(A, dev(psi), 0.7)
    assert np.array_equal(A.cpu().numpy(), oracle.brox_set_matrix(A0, psi, 0.7))
    A4 = dev(A0[:4])
    diag = K.brox_set_redblack_matrix(A4, dev(psi), 0.7)
    R4, Rd = oracle.brox_set_matrix(A0[:4], psi, 0.7, redblack=True)
    assert np.array_equal(A4.cpu().numpy(), R4) and np.array_equal(diag.cpu().numpy(), Rd)


@pytest.mark.parametrize("shape", [(1, 1), (37, 53)])
def test_brox_precondition_bitexact(K, shape):
    rng = np.random.default_rng(11)
    H, W = shape
    A = rng.normal(size=(5, H, W)).astype(np.float32)
    d0, d1, od, x0, x1 = (rng.normal(size=(H, W)).astype(np.float32) for _ in range(5))
    d0.flat[::5] = 0.0
    d1.flat[::5] = 0.0
    od.flat[::5] = 0.0
    A[2].flat[::5] = 0.0          # det = 0: the pass-through branch
    y0, y1 = K.brox_precondition(dev(A), dev(d0), dev(d1), dev(od), dev(x0), dev(x1))
    r0, r1 = oracle.brox_precondition(A, d0, d1, od, x0, x1)
    assert np.array_equal(y0.cpu().numpy(), r0) and np.array_equal(y1.cpu().numpy(), r1)


def test_hm_red_black_sweep_bitexact():
    os.environ["HM_GPU0"] = "1"
    from hindemith_amd.hindemith import SpMat2D, defines, hm
    saved = dict(defines)
    defines.update(nd20=1, nd21=1, nt20=8, nt21=8, ne20=2, ne21=2)
    try:
        @hm
        def sweep(A, psi, alpha, du, b0, diag0, diag1, offdiag):
            set_brox_matrix(A, psi, alpha)  # noqa: F821
            for i in range(3):
                r0 = b0 - A * du
                z0, z1 = applyBroxPreconditioner(A, diag0, diag1, offdiag, r0, r0)  # noqa: F821
                du = du + red(z0)  # noqa: F821
                r0 = b0 - A * du
                z0, z1 = applyBroxPreconditioner(A, diag0, diag1, offdiag, r0, r0)  # noqa: F821
                du = du + black(z0)  # noqa: F821
            return du

        rng = np.random.default_rng(5)
        H, W = 48, 70
        psi, du, b0, d0, d1 = (rng.uniform(0.5, 1.5, (H, W)).astype(np.float32)
                               for _ in range(5))
        od = (0.1 * rng.normal(size=(H, W))).astype(np.float32)
        A0 = np.zeros((5, H, W), np.float32)
        got = sweep(A=SpMat2D(A0.copy(), OFFX, OFFY), psi=psi, alpha=np.float32(0.5),
                    du=du.copy(), b0=b0, diag0=d0, diag1=d1, offdiag=od)
        A = oracle.brox_set_matrix(A0, psi, 0.5)
        u = du.copy()
        for _ in range(3):
            for mask in (oracle.red, oracle.black):
                r0 = (b0 - oracle.dia2d(A, OFFX, OFFY, u)).astype(np.float32)
                z0, _ = oracle.brox_precondition(A, d0, d1, od, r0, r0)
                u = (u + mask(z0)).astype(np.float32)
        assert np.array_equal(got, u)
    finally:
        defines.clear()
        defines.update(saved)
