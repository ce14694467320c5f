"""Pins of the oracle's Brox SOR building blocks.

SetBroxMatrix (SetBroxMatrix.cpp:34-79) is pinned by the operator its planes
define under the SpMat2D convention (coefficient at the neighbour's index,
assembled by the independent numpy DIA->CSR conversion): it must equal the
graph Laplacian built here from edge lists with weight alpha*psi at each edge's
upper / left pixel -- symmetric, zero row sums.  The red-black form
(SetBroxRedblackMatrix.cpp:37-81) must carry the same off-diagonals and
diagonal.  ApplyBroxPreconditioner (ApplyBroxPreconditioner.cpp:46-72) is
pinned by the block systems it solves and its 1e-4 threshold; red/black by the
checkerboard they split x into.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle

OFFX = np.array([0, -1, 0, 1, 0], np.int32)
OFFY = np.array([-1, 0, 0, 0, 1], np.int32)


def edge_laplacian(psi, alpha):
    H, W = psi.shape
    idx = np.arange(H * W).reshape(H, W)
    rows, cols, vals = [], [], []
    # horizontal edges (p, p+1) weighted at the left pixel, vertical (p, p+W) at the upper
    for a, b, w in ((idx[:, :-1], idx[:, 1:], psi[:, :-1]), (idx[:-1, :], idx[1:, :], psi[:-1, :])):
        a, b, w = a.ravel(), b.ravel(), alpha * w.ravel().astype(np.float64)
        rows += [a, b, a, b]
        cols += [b, a, a, b]
        vals += [-w, -w, w, w]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(H * W, H * W))


@pytest.mark.parametrize("shape", [(1, 1), (1, 6), (5, 1), (7, 9)])
def test_brox_matrix_is_the_psi_weighted_laplacian(shape):
    rng = np.random.default_rng(1)
    H, W = shape
    psi = rng.uniform(0.2, 2.0, (H, W)).astype(np.float32)
    alpha = 0.75
    A0 = np.full((5, H, W), np.nan, np.float32)   # unwritten entries must never be read
    A = oracle.brox_set_matrix(A0, psi, alpha)
    M = oracle.np_dia2d_csr(A, OFFX, OFFY)
    assert np.isfinite(M.data).all()
    L = edge_laplacian(psi, alpha)
    assert np.allclose(M.toarray(), L.toarray(), rtol=1e-6, atol=1e-6)
    assert np.allclose(M.toarray(), M.toarray().T, rtol=0, atol=1e-6)
    assert np.allclose(np.asarray(M.sum(axis=1)).ravel(), 0.0, atol=1e-5)


def test_redblack_form_carries_the_same_matrix():
    rng = np.random.default_rng(2)
    psi = rng.uniform(0.2, 2.0, (6, 8)).astype(np.float32)
    A = oracle.brox_set_matrix(np.zeros((5, 6, 8), np.float32), psi, 1.3)
    A4, diag = oracle.brox_set_matrix(np.zeros((4, 6, 8), np.float32), psi, 1.3, redblack=True)
    assert np.array_equal(A4, A[[0, 1, 3, 4]])
    assert np.array_equal(diag, A[2])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_brox_precondition_solves_the_block_system(dtype):
    rng = np.random.default_rng(3)
    H, W = 9, 11
    A = rng.uniform(-1, 1, (5, H, W))
    diag0 = rng.uniform(2, 3, (H, W))
    diag1 = rng.uniform(2, 3, (H, W))
    off = rng.uniform(-0.5, 0.5, (H, W))
    x0, x1 = rng.normal(size=(2, H, W))
    y0, y1 = oracle.brox_precondition(A, diag0, diag1, off, x0, x1, dtype)
    a, d = diag0 + A[2], diag1 + A[2]
    tol = 1e-12 if dtype == np.float64 else 1e-5
    assert np.allclose(a * y0 + off * y1, x0, rtol=tol, atol=tol)
    assert np.allclose(off * y0 + d * y1, x1, rtol=tol, atol=tol)


def test_brox_precondition_threshold_is_1e_4():
    A = np.zeros((3, 1, 2))
    a = np.array([[1.0, 1.0]])
    d = np.array([[1e-3, 1e-5]])          # det = 1e-3 (solved), 1e-5 (passed through)
    off = np.zeros((1, 2))
    y0, _ = oracle.brox_precondition(A, a, d, off, np.ones((1, 2)), np.ones((1, 2)), np.float64)
    y0f, _ = oracle.brox_precondition(A, a, d, off, np.ones((1, 2)), np.ones((1, 2)))
    assert y0[0, 0] == pytest.approx(1.0) and y0[0, 1] == 1.0
    assert y0f[0, 1] == 1.0


def test_red_black_checkerboard():
    x = np.arange(1, 1 + 5 * 7, dtype=np.float32).reshape(5, 7)
    r, b = oracle.red(x), oracle.black(x)
    assert np.array_equal(r + b, x)
    assert r[0, 0] == x[0, 0] and b[0, 0] == 0 and r[0, 1] == 0 and b[0, 1] == x[0, 1]
    assert np.count_nonzero(r * b) == 0
