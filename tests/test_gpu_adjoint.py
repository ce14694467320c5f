"""Adjoint Horn-Schunck on the GPU (autograd Function over hmx_hs_adjoint).

Checks: (1) against the fp64 reverse-mode oracle on the CPU; (2) against
fp64 central finite differences (eps = 1e-3) computed with the GPU fp64
forward on 1024 random pixels of the 8 x 512 x 512 workload; (3) the
checkpoint interval does not change the result (recomputed states are
bit-identical).  Tolerance: 1e-3 relative (norm over the sampled entries).
"""
import numpy as np
import pytest

import oracle
from hindemith_amd import synthetic as S

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
TAPS = S.hs_stencil_taps()


def batch(B, H, W, seed):
    rng = np.random.default_rng(seed)
    I1, I2, fx, fy = [], [], [], []
    for b in range(B):
        tx, ty = rng.uniform(-3, 3, 2)
        a, c, x, y = S.hs_pair(H, W, seed=seed * 100 + b, flow="translation", trans=(tx, ty))
        I1.append(a), I2.append(c), fx.append(x), fy.append(y)
    alpha = rng.uniform(0.5, 2.0, B).astype(np.float32)
    return (np.stack(I1), np.stack(I2), np.stack(fx), np.stack(fy), alpha)


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def run_grad(I1, I2, fx, fy, alpha, iters, ckpt):
    from hindemith_amd.autograd import horn_schunck
    d = "cuda"
    tI1 = torch.tensor(I1, device=d, requires_grad=True)
    tI2 = torch.tensor(I2, device=d, requires_grad=True)
    ta = torch.tensor(alpha, device=d, requires_grad=True)
    u, v = horn_schunck(tI1, tI2, ta, iters=iters, ckpt_every=ckpt)
    tfx, tfy = torch.tensor(fx, device=d), torch.tensor(fy, device=d)
    loss = ((u - tfx) ** 2 + (v - tfy) ** 2).mean()
    loss.backward()
    torch.cuda.synchronize()
    return (tI1.grad.cpu().numpy(), tI2.grad.cpu().numpy(), ta.grad.cpu().numpy(),
            u.detach().cpu().numpy(), v.detach().cpu().numpy())


def test_forward_matches_oracle_bitexact():
    I1, I2, fx, fy, alpha = batch(3, 45, 70, 1)
    from hindemith_amd.autograd import horn_schunck
    u, v = horn_schunck(torch.tensor(I1, device="cuda"), torch.tensor(I2, device="cuda"),
                        torch.tensor(alpha, device="cuda"), iters=33, ckpt_every=7)
    u, v = u.cpu().numpy(), v.cpu().numpy()
    for b in range(3):
        a2 = np.float32(alpha[b] * alpha[b])
        ru, rv = oracle.hs_solve(I1[b], I2[b], a2, TAPS, 33)
        assert np.array_equal(u[b], ru) and np.array_equal(v[b], rv)


def test_small_vs_fp64_oracle_gradient():
    B, H, W, n = 2, 40, 52, 30
    I1, I2, fx, fy, alpha = batch(B, H, W, 2)
    gI1, gI2, ga, u, v = run_grad(I1, I2, fx, fy, alpha, n, 7)
    scale = 2.0 / (B * H * W)
    for b in range(B):
        a64 = float(alpha[b])
        Ix, Iy, It = oracle.hs_setup(I1[b], I2[b], TAPS["Dx"], TAPS["Dy"], TAPS["Dt"], np.float64)
        u64, v64 = oracle.hs_jacobi(Ix, Iy, It, a64 * a64, TAPS["Avg"], n, dtype=np.float64)
        r1, r2, ra, _, _ = oracle.hs_grad(I1[b], I2[b], a64, TAPS, n, scale * (u64 - fx[b]),
                                          scale * (v64 - fy[b]))
        assert relerr(gI1[b], r1) < 1e-3 and relerr(gI2[b], r2) < 1e-3
        assert abs(ga[b] - ra) <= 1e-3 * abs(ra)


@pytest.mark.parametrize("shape", [(1, 9), (9, 1), (2, 3), (3, 70), (33, 65)])
def test_border_heavy_shapes_vs_fp64_oracle_gradient(shape):
    """Images where most pixels lie on a border (a pixel on both the left and the
    right border when W = 1, on top and bottom when H <= 2): the reverse sweep's
    transposed clamp comes from its ghost gathers.  12 steps in checkpoint
    segments of 5, so reverse launches of fewer than 8 steps are covered too."""
    H, W = shape
    B, n = 2, 12
    I1, I2, fx, fy, alpha = batch(B, H, W, 4)
    gI1, gI2, ga, u, v = run_grad(I1, I2, fx, fy, alpha, n, 5)
    scale = 2.0 / (B * H * W)
    for b in range(B):
        a64 = float(alpha[b])
        Ix, Iy, It = oracle.hs_setup(I1[b], I2[b], TAPS["Dx"], TAPS["Dy"], TAPS["Dt"], np.float64)
        u64, v64 = oracle.hs_jacobi(Ix, Iy, It, a64 * a64, TAPS["Avg"], n, dtype=np.float64)
        r1, r2, ra, _, _ = oracle.hs_grad(I1[b], I2[b], a64, TAPS, n, scale * (u64 - fx[b]),
                                          scale * (v64 - fy[b]))
        assert relerr(gI1[b], r1) < 1e-3 and relerr(gI2[b], r2) < 1e-3, (shape, b)
        assert abs(ga[b] - ra) <= 1e-3 * abs(ra), (shape, b)


def test_checkpoint_interval_invariance():
    I1, I2, fx, fy, alpha = batch(2, 33, 47, 3)
    outs = [run_grad(I1, I2, fx, fy, alpha, 25, c) for c in (1, 6, 25, 40)]
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", [(100, 200), (37, 66)])
def test_ring_build_matches_generic_build_bitwise(shape, monkeypatch):
    """The reverse sweep's build for the canonical 8-neighbour stencil (every
    tap a constant LDS offset) and its generic build (taps from the table)
    perform the same arithmetic in the same order: identical results.  At
    100 x 200 one tile runs the unclipped loop, the rest the clipped one; 21
    steps in segments of 10 cover reverse launches of 8, 2 and 1 steps."""
    H, W = shape
    I1, I2, fx, fy, alpha = batch(2, H, W, 6)
    ring = run_grad(I1, I2, fx, fy, alpha, 21, 10)
    monkeypatch.setenv("HMX_ADJ_GENERIC", "1")
    generic = run_grad(I1, I2, fx, fy, alpha, 21, 10)
    for a, b in zip(ring, generic):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", [(100, 200), (37, 66)])
def test_register_linearisation_matches_staged_bitwise(shape, monkeypatch):
    """The segment recompute's one-barrier form (each step's (ubar, vbar)
    stored from registers) and its two-barrier form (staged through LDS, stored
    in tile order) recompute the same states and store the same values:
    identical gradients.  21 steps in segments of 10 cover recompute launches
    of 8, 2 and 1 steps."""
    H, W = shape
    I1, I2, fx, fy, alpha = batch(2, H, W, 7)
    reg = run_grad(I1, I2, fx, fy, alpha, 21, 10)
    monkeypatch.setenv("HMX_LIN_STAGED", "1")
    staged = run_grad(I1, I2, fx, fy, alpha, 21, 10)
    for a, b in zip(reg, staged):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape,iters,ckpt", [((100, 200), 37, 10), ((37, 66), 21, 7),
                                              ((40, 52), 8, 8)])
def test_kept_linearisation_and_planes_match_recompute_bitwise(shape, iters, ckpt):
    """The backward as autograd runs it (the forward's last-segment
    linearisation and derivative planes reused) against the C ABI called with
    neither (every segment recomputed, the planes rebuilt): the same gradients
    bit for bit, for last segments of 7, 7 and 8 iterations."""
    import ctypes
    from hindemith_amd import autograd as AG
    from hindemith_amd._lib import check, lib
    from hindemith_amd.kernels import _ptr, stream_ptr
    H, W = shape
    B = 2
    I1, I2, fx, fy, alpha = batch(B, H, W, 8)
    d = "cuda"
    tI1 = torch.tensor(I1, device=d, requires_grad=True)
    tI2 = torch.tensor(I2, device=d, requires_grad=True)
    ta = torch.tensor(alpha, device=d, requires_grad=True)
    u, v = AG.horn_schunck(tI1, tI2, ta, iters=iters, ckpt_every=ckpt)
    rng = np.random.default_rng(1)
    gu = torch.tensor(rng.normal(size=(B, H, W)).astype(np.float32), device=d)
    gv = torch.tensor(rng.normal(size=(B, H, W)).astype(np.float32), device=d)
    g1, g2, ga = torch.autograd.grad((u, v), (tI1, tI2, ta), (gu, gv))

    L = lib()
    keep, targs = AG._tap_args(AG.default_taps())
    null = ctypes.c_void_p(0)
    dI1, dI2, dal = (t.detach().contiguous() for t in (tI1, tI2, ta))
    uu, vv = torch.zeros_like(dI1), torch.zeros_like(dI1)
    nck = L.hmx_hs_ckpt_count(iters, ckpt)
    ck = torch.empty((nck, 2, B, H, W), device=d)
    fws = torch.empty(int(L.hmx_hs_forward_workspace(B, H, W)) // 4 + 1, device=d)
    check(L.hmx_hs_forward_ckpt(_ptr(dI1), _ptr(dI2), B, H, W, _ptr(dal), *targs, iters, ckpt,
                                _ptr(uu), _ptr(vv), _ptr(ck), null, _ptr(fws), stream_ptr()),
          "forward")
    assert torch.equal(uu, u.detach()) and torch.equal(vv, v.detach())
    r1, r2, rga = torch.empty_like(dI1), torch.empty_like(dI1), torch.empty_like(dal)
    ws = torch.empty(int(L.hmx_hs_adjoint_workspace(B, H, W, ckpt)) // 4 + 1, device=d)
    check(L.hmx_hs_adjoint(_ptr(dI1), _ptr(dI2), B, H, W, _ptr(dal), *targs, iters, ckpt,
                           _ptr(ck), null, null, _ptr(gu), _ptr(gv), _ptr(r1), _ptr(r2),
                           _ptr(rga), null, null, _ptr(ws), stream_ptr()), "adjoint")
    torch.cuda.synchronize()
    assert torch.equal(g1, r1) and torch.equal(g2, r2) and torch.equal(ga, rga)


@pytest.fixture(scope="module")
def config5():
    """8 pairs at 512x512, 200 iterations, checkpoints every 20."""
    I1, I2, fx, fy, alpha = batch(8, 512, 512, 5)
    grads = run_grad(I1, I2, fx, fy, alpha, 200, 20)
    return I1, I2, fx, fy, alpha, grads


def test_config5_vs_fp64_oracle_on_1024_pixels(config5):
    I1, I2, fx, fy, alpha, (gI1, gI2, ga, u, v) = config5
    B, H, W = I1.shape
    rng = np.random.default_rng(0)
    scale = 2.0 / (B * H * W)
    got, ref = [], []
    for b in (0, 5):  # the fp64 CPU reference stores every state: two pairs keep it quick
        a64 = float(alpha[b])
        Ix, Iy, It = oracle.hs_setup(I1[b], I2[b], TAPS["Dx"], TAPS["Dy"], TAPS["Dt"], np.float64)
        u64, v64 = oracle.hs_jacobi(Ix, Iy, It, a64 * a64, TAPS["Avg"], 200, dtype=np.float64)
        r1, r2, ra, _, _ = oracle.hs_grad(I1[b], I2[b], a64, TAPS, 200, scale * (u64 - fx[b]),
                                          scale * (v64 - fy[b]))
        idx = rng.integers(0, H * W, 512)
        got += [gI1[b].reshape(-1)[idx], gI2[b].reshape(-1)[idx]]
        ref += [r1.reshape(-1)[idx], r2.reshape(-1)[idx]]
        assert abs(ga[b] - ra) <= 1e-3 * abs(ra)
    err = relerr(np.concatenate(got), np.concatenate(ref))
    print(f"config5 adjoint vs fp64 oracle, 1024 pixels: rel err {err:.2e}")
    assert err < 1e-3


def test_config5_vs_central_fd(config5):
    from hindemith_amd.autograd import hs_solve_f64
    I1, I2, fx, fy, alpha, (gI1, gI2, ga, u, v) = config5
    B, H, W = I1.shape
    eps = 1e-3
    d = "cuda"
    J1 = torch.tensor(I1, device=d, dtype=torch.float64)
    J2 = torch.tensor(I2, device=d, dtype=torch.float64)
    A = torch.tensor(alpha, device=d, dtype=torch.float64)
    tfx = torch.tensor(fx, device=d, dtype=torch.float64)
    tfy = torch.tensor(fy, device=d, dtype=torch.float64)

    def per_pair_loss(j1, j2, a):
        uu, vv = hs_solve_f64(j1, j2, a, 200)
        return (((uu - tfx) ** 2 + (vv - tfy) ** 2).sum(dim=(1, 2)) / (B * H * W)).cpu().numpy()

    rng = np.random.default_rng(1)
    got, fd = [], []
    n_rounds = 64  # one pixel per pair per round and image: 64 * 8 * 2 = 1024 pixels
    for r in range(n_rounds):
        for img in (0, 1):
            pix = rng.integers(0, H * W, B)
            rows, cols = pix // W, pix % W
            base = J1 if img == 0 else J2
            plus, minus = base.clone(), base.clone()
            bi = torch.arange(B, device=d)
            plus[bi, torch.tensor(rows, device=d), torch.tensor(cols, device=d)] += eps
            minus[bi, torch.tensor(rows, device=d), torch.tensor(cols, device=d)] -= eps
            if img == 0:
                lp, lm = per_pair_loss(plus, J2, A), per_pair_loss(minus, J2, A)
            else:
                lp, lm = per_pair_loss(J1, plus, A), per_pair_loss(J1, minus, A)
            fd.append((lp - lm) / (2 * eps))
            g = gI1 if img == 0 else gI2
            got.append(g[np.arange(B), rows, cols])
    err = relerr(np.concatenate(got), np.concatenate(fd))
    lp = per_pair_loss(J1, J2, A + eps)
    lm = per_pair_loss(J1, J2, A - eps)
    fda = (lp - lm) / (2 * eps)
    aerr = relerr(ga, fda)
    print(f"config5 adjoint vs central FD (eps=1e-3), 1024 pixels: rel err {err:.2e}; "
          f"alpha rel err {aerr:.2e}")
    assert err < 1e-3 and aerr < 1e-3
