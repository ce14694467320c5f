"""Pins of the oracle's preconditioned-CG Horn-Schunck operations.

ApplyHSPreconditioner (ApplyHSPreconditioner.cpp:46-97) is pinned by the
system it solves: the 2x2 blocks are assembled here from an independent
neighbour count (scipy.ndimage), and M y must give back x.  HSMatrixPowers
(HSMatrixPowers.cpp:72-126) is pinned against the operator assembled as a
scipy.sparse matrix from Kronecker sums of 1-D Neumann Laplacians, by the
operator's symmetry, and by the constant-field null space.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.sparse as sp

import oracle

CROSS = np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]], float)


def neighbours(H, W):
    """In-image 4-neighbour count per pixel."""
    return ndi.correlate(np.ones((H, W)), CROSS, mode="constant", cval=0.0)


def blocks(H, W, rng, s2):
    diag0 = rng.uniform(1.0, 2.0, (H, W))
    diag1 = rng.uniform(1.0, 2.0, (H, W))
    off = rng.uniform(-0.5, 0.5, (H, W))
    x0, x1 = rng.normal(size=(2, H, W))
    return diag0, diag1, off, x0, x1


@pytest.mark.parametrize("shape", [(1, 1), (1, 9), (7, 1), (13, 17)])
@pytest.mark.parametrize("s2", [0.25, -0.1])
def test_precond_solves_the_block_system(shape, s2):
    rng = np.random.default_rng(1)
    H, W = shape
    diag0, diag1, off, x0, x1 = blocks(H, W, rng, s2)
    y0, y1 = oracle.hs_precondition(s2, diag0, diag1, off, x0, x1, np.float64)
    n = neighbours(H, W)
    a, d = diag0 + n * s2, diag1 + n * s2
    assert np.allclose(a * y0 + off * y1, x0, rtol=1e-12, atol=1e-12)
    assert np.allclose(off * y0 + d * y1, x1, rtol=1e-12, atol=1e-12)
    f0, f1 = oracle.hs_precondition(s2, diag0, diag1, off, x0, x1)
    assert np.allclose(f0, y0, rtol=1e-5, atol=1e-5) and np.allclose(f1, y1, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_precond_singular_blocks_pass_through(dtype):
    # a = d = b = 2 exactly (diag = 2 - n*s2 is exact for s2 = 0.25): det = 0
    H, W, s2 = 5, 6, 0.25
    n = neighbours(H, W)
    diag = 2.0 - n * s2
    off = np.full((H, W), 2.0)
    x0 = np.arange(H * W, dtype=float).reshape(H, W)
    x1 = -x0
    y0, y1 = oracle.hs_precondition(s2, diag, diag, off, x0, x1, dtype)
    assert np.array_equal(y0, x0.astype(dtype)) and np.array_equal(y1, x1.astype(dtype))


def test_precond_threshold_is_1e_15():
    # |det| = 1e-14 is solved, |det| = 1e-16 passes through
    s2 = 0.0
    a = np.array([[1.0, 1.0]])
    off = np.array([[np.sqrt(1.0 - 1e-14), 0.0]])
    d = np.array([[1.0, 1e-16]])
    y0, _ = oracle.hs_precondition(s2, a, d, off, np.ones((1, 2)), np.zeros((1, 2)), np.float64)
    assert abs(y0[0, 0] - 1.0) > 1.0   # solved: 1 / det ~ 1e14
    assert y0[0, 1] == 1.0              # det 1e-16: passed through


def neumann_laplacian(n):
    """Path-graph Laplacian: degree on the diagonal, -1 to each neighbour."""
    if n == 1:
        return sp.csr_matrix((1, 1))
    main = np.full(n, 2.0)
    main[0] = main[-1] = 1.0
    return sp.diags([main, -np.ones(n - 1), -np.ones(n - 1)], [0, -1, 1], format="csr")


def hs_operator(lam, Ix, Iy):
    H, W = Ix.shape
    L = sp.kron(sp.identity(H), neumann_laplacian(W)) + sp.kron(neumann_laplacian(H),
                                                                 sp.identity(W))
    Dx, Dy = sp.diags(Ix.ravel()), sp.diags(Iy.ravel())
    return sp.bmat([[Dx @ Dx + lam * L, Dx @ Dy], [Dy @ Dx, Dy @ Dy + lam * L]], format="csr")


@pytest.mark.parametrize("shape", [(1, 1), (1, 8), (6, 1), (9, 11)])
def test_powers_equal_sparse_operator_powers(shape):
    rng = np.random.default_rng(2)
    H, W = shape
    Ix, Iy, r0, r1 = rng.normal(size=(4, H, W))
    lam, steps = 0.7, 6
    V0, V1 = oracle.hs_matrix_powers(lam, Ix, Iy, r0, r1, steps, np.float64)
    A = hs_operator(lam, Ix, Iy)
    x = np.concatenate([r0.ravel(), r1.ravel()])
    for s in range(steps):
        got = np.concatenate([V0[s].ravel(), V1[s].ravel()])
        assert np.allclose(got, x, rtol=1e-11, atol=1e-11 * np.abs(x).max()), s
        x = A @ x


def test_powers_operator_is_symmetric():
    rng = np.random.default_rng(3)
    H, W = 7, 10
    Ix, Iy, a0, a1, b0, b1 = rng.normal(size=(6, H, W))
    lam = 1.3
    Aa = oracle.hs_matrix_powers(lam, Ix, Iy, a0, a1, 2, np.float64)
    Ab = oracle.hs_matrix_powers(lam, Ix, Iy, b0, b1, 2, np.float64)
    lhs = np.sum(a0 * Ab[0][1]) + np.sum(a1 * Ab[1][1])
    rhs = np.sum(Aa[0][1] * b0) + np.sum(Aa[1][1] * b1)
    assert lhs == pytest.approx(rhs, rel=1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_powers_constant_field_is_in_the_laplacian_null_space(dtype):
    H, W = 5, 8
    z = np.zeros((H, W))
    c = np.full((H, W), 3.5)
    V0, V1 = oracle.hs_matrix_powers(0.9, z, z, c, -c, 4, dtype)
    assert np.array_equal(V0[0], c.astype(dtype))
    if dtype == np.float64:
        assert np.all(V0[1:] == 0) and np.all(V1[1:] == 0)
    else:
        # the fp32 form contracts "U += -lam*u_nb; U += lam*u" into two fmas, which
        # leaves the rounding error of one product per neighbour (< 1 ulp of lam*c)
        bound = 4 * np.spacing(np.float32(0.9 * 3.5))
        assert np.max(np.abs(V0[1])) <= bound and np.max(np.abs(V1[1])) <= bound


def test_powers_fp32_close_to_fp64():
    rng = np.random.default_rng(4)
    Ix, Iy, r0, r1 = rng.normal(size=(4, 16, 21))
    f = oracle.hs_matrix_powers(0.5, Ix, Iy, r0, r1, 5)
    d = oracle.hs_matrix_powers(0.5, Ix, Iy, r0, r1, 5, np.float64)
    for s in range(5):
        scale = np.abs(d[0][s]).max()
        assert np.max(np.abs(f[0][s] - d[0][s])) <= 1e-5 * scale
