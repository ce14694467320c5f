"""Batched persistent Horn-Schunck (hmx_hs_batched) on the GPU.

Per-pair results must equal the single-pair path and the fp32 oracle bit for
bit (same tile body, same per-op rounding); graph replays, interleaved with
other GPU work, must all agree; a host-forced tiny barrier spin limit must
end in HMX_ERR_TIMEOUT with the grid drained (later calls still work).
"""
import numpy as np
import pytest

import oracle
from hindemith_amd import synthetic as S

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
TAPS = S.hs_stencil_taps()


def make(n, seed, side=(20, 300), iters=23):
    from hindemith_amd.batched import BatchedHS
    pairs = S.batched_pairs(n, seed=seed, side=side)
    B = BatchedHS([torch.from_numpy(p["I1"]).cuda() for p in pairs],
                  [torch.from_numpy(p["I2"]).cuda() for p in pairs],
                  [p["alpha"] for p in pairs], iters=iters)
    return pairs, B


def check_oracle(pairs, out, iters):
    for p, (u, v) in zip(pairs, out):
        a2 = np.float32(p["alpha"]) * np.float32(p["alpha"])
        ru, rv = oracle.hs_solve(p["I1"], p["I2"], a2, TAPS, iters)
        assert np.array_equal(u.cpu().numpy(), ru) and np.array_equal(v.cpu().numpy(), rv)


@pytest.mark.parametrize("graph", [False, True])
def test_batched_matches_oracle(graph):
    pairs, B = make(9, seed=1)
    out = B.run(graph=graph, sync=True)
    torch.cuda.synchronize()
    check_oracle(pairs, out, 23)


def test_all_variants_and_ks():
    from hindemith_amd._lib import lib
    pairs, B = make(5, seed=2, iters=17)
    L = lib()
    for vi in range(L.hmx_hs_batched_variant_count()):
        kmax = ctypes_info(vi)
        for k in sorted({1, 2, kmax}):
            out = B.run(graph=False, sync=True, variant=vi, k=k)
            torch.cuda.synchronize()
            check_oracle(pairs, out, 17)


def test_phase_counts_one_even_odd():
    """Every variant at phase counts 1 (the setup's copy), even (two plane
    pairs) and odd above 1 (the stepped form's third plane pair; the
    persistent form's copy), with a short last phase, against the oracle."""
    from hindemith_amd._lib import lib
    for iters in (1, 8, 13, 16, 24, 35):
        pairs, B = make(4, seed=12, side=(30, 200), iters=iters)
        refs = [oracle.hs_solve(p["I1"], p["I2"],
                                np.float32(p["alpha"]) * np.float32(p["alpha"]), TAPS, iters)
                for p in pairs]
        for vi in range(lib().hmx_hs_batched_variant_count()):
            k = 8 if ctypes_info(vi) >= 8 else 4
            for graph in (False, True):
                out = B.run(graph=graph, sync=True, variant=vi, k=k)
                torch.cuda.synchronize()
                for (ru, rv), (u, v) in zip(refs, out):
                    assert np.array_equal(u.cpu().numpy(), ru), (iters, vi, k, graph)
                    assert np.array_equal(v.cpu().numpy(), rv), (iters, vi, k, graph)


def ctypes_info(vi):
    import ctypes
    from hindemith_amd._lib import lib
    km = ctypes.c_int()
    lib().hmx_hs_batched_variant_info(vi, None, None, None, ctypes.byref(km))
    return km.value


def test_graph_replays_consistent_with_interleaved_work():
    pairs, B = make(7, seed=3, iters=30)
    ref = [(u.clone(), v.clone()) for u, v in B.run(graph=False, sync=True)]
    x = torch.randn(1 << 20, device="cuda")
    for rep in range(4):
        out = B.run(graph=True)
        y = (x * 2.0).sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        assert float(y) == float(y)
        for (u, v), (ru, rv) in zip(out, ref):
            assert torch.equal(u, ru) and torch.equal(v, rv), rep


def test_tune_then_run_matches():
    pairs, B = make(6, seed=4, iters=20)
    vi, k = B.tune()
    out = B.run(graph=True, sync=True)
    torch.cuda.synchronize()
    check_oracle(pairs, out, 20)


def test_forced_barrier_timeout_drains_and_recovers():
    """The persistent form's bounded barrier wait (variant 1; the stepped
    variants have no grid barrier)."""
    from hindemith_amd._lib import HindemithError
    pairs, B = make(12, seed=5, side=(200, 400), iters=16)
    with pytest.raises(HindemithError, match="gave up"):
        B.run(graph=False, sync=True, spin_limit=1, variant=1, k=8)
    out = B.run(graph=False, sync=True)
    torch.cuda.synchronize()
    check_oracle(pairs, out, 16)


def test_forced_barrier_timeout_reported_without_sync():
    """A persistent call reports a barrier give-up even when the caller did not
    ask for a synchronising call (the result would be partial)."""
    from hindemith_amd._lib import HindemithError
    pairs, B = make(12, seed=5, side=(200, 400), iters=16)
    with pytest.raises(HindemithError, match="gave up"):
        B.run(graph=False, sync=False, spin_limit=1, variant=1, k=8)
    with pytest.raises(HindemithError, match="gave up"):
        B.run(graph=True, sync=False, spin_limit=1, variant=1, k=8)
    out = B.run(graph=False, sync=False, variant=1, k=8)
    torch.cuda.synchronize()
    check_oracle(pairs, out, 16)


def test_graph_call_on_a_plan_made_eagerly():
    """A signature first run eagerly (as the tuner does) is captured when a
    graph call first asks for it; every replay equals the oracle."""
    pairs, B = make(5, seed=8, iters=21)
    B.run(graph=False, sync=True)
    for _ in range(3):
        out = B.run(graph=True)
        torch.cuda.synchronize()
        check_oracle(pairs, out, 21)


def test_batched_equals_per_pair_plain_solves():
    pairs, B = make(8, seed=6, iters=25)
    batched = [(u.clone(), v.clone()) for u, v in B.run(graph=True, sync=True)]
    plain = B.solve_each()
    torch.cuda.synchronize()
    for (bu, bv), (pu, pv) in zip(batched, plain):
        assert torch.equal(bu, pu) and torch.equal(bv, pv)


def test_full_size_batch_every_variant_stays_resident():
    """64 pairs with sides U(64, 1024): every variant's grid is the full
    resident size, so a grid larger than what the GPU holds at once would end
    in a barrier time-out here.  Results must equal the per-pair plain path."""
    from hindemith_amd.batched import batched_variants
    pairs, B = make(64, seed=0, side=(64, 1024), iters=9)
    plain = [(u.clone(), v.clone()) for u, v in B.solve_each()]
    for vi, info in enumerate(batched_variants()):
        assert 1 <= min(info["api_per_cu"], info["own_per_cu"]), info
        for k in sorted({2, info["kmax"]}):
            out = B.run(graph=False, sync=True, spin_limit=1 << 22, variant=vi, k=k)
            torch.cuda.synchronize()
            for i, ((bu, bv), (pu, pv)) in enumerate(zip(out, plain)):
                assert torch.equal(bu, pu) and torch.equal(bv, pv), (vi, k, i)
