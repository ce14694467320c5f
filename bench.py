#!/usr/bin/env python3
"""Horn-Schunck throughput benchmark on one MI355X.

Workload: 1920x1080 fp32 pair (32 Gaussian blobs + N(0, 0.01) noise; frame 2
= frame 1 warped by the affine flow rotation 0.5 deg, scale 1.01, translation
(2, -1) px), alpha = 1, 500 Jacobi iterations from u = v = 0, replicated
borders.  One step = one full solve (derivative setup + 500 fused Jacobi
iterations) with inputs already resident in HBM.

Prints one JSON line: value = Mpixel*iter/s, plus
  roofline: the fused Jacobi kernel against the resource that binds it, the
            vector ALU: issued VALU lane-instructions per launch (rocprofv3
            SQ_INSTS_VALU of the committed counter summary) over the mean
            launch time, timed with HIP events on the launch stream, against
            the non-packed fp32 issue peak; with the same launches against HBM
            (28 B per pixel per iteration of algorithmic bytes, and the
            measured FETCH/WRITE traffic) and against LDS;
  cpu_baseline: the oracle's C/OpenMP restatement on this host's cores, on a
            bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PX_ITER = 28
# Non-packed fp32 VALU issue peak of MI355X: 256 CUs x 4 SIMDs x 16 lanes x
# 2.4 GHz lane-instructions per second (a wave64 instruction occupies a SIMD
# for 4 cycles).  Packed fp32 is disabled in these builds.
VALU_LANE_PEAK = 256 * 4 * 16 * 2.4e9
# LDS: 256 CUs x 128 B/clk x 2.4 GHz.
LDS_PEAK_BPS = 256 * 128 * 2.4e9
COUNTERS = os.path.join(HERE, "profiles", "hs_jacobi_1920x1080_counters.json")


def parse():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    # 300 timed steps (about 0.6 s at 1080p) after 50 untimed ones: the first
    # tens of milliseconds of GPU work run slower than the sustained rate
    # (10 steps after 2 measured 477 k against 509 k Mpx*iter/s on one box)
    p.add_argument("--steps", type=int, default=300)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--iters", type=int, default=500)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-tune", action="store_true", help="use the default kernel choice")
    p.add_argument("--variant", type=int, default=None)
    p.add_argument("--k", type=int, default=None)
    p.add_argument("--no-graph", action="store_true",
                   help="launch each step eagerly instead of replaying a captured hipGraph")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the secondary workloads (4K, bandwidth kernels, batched, adjoint)")
    p.add_argument("--copy-mb", type=int, default=0,
                   help="only run a plain device copy of this many MiB (counter calibration)")
    p.add_argument("--secondary",
                   choices=["4k", "bandwidth", "batched", "adjoint", "frontend", "hscg"],
                   default=None, help="only run this secondary workload and print its JSON")
    p.add_argument("--traffic-sweep", action="store_true",
                   help="only launch every (variant, k) of the fused Jacobi kernel 3 times, in "
                        "the printed order (for per-configuration rocprofv3 counters)")
    return p.parse_args()


def copy_calibration(mb, reps=5):
    """A plain device-to-device copy of known size, for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = mb * (1 << 20) // 4
    x = torch.rand(n, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "reps": reps}))


def traffic_sweep(H, W):
    """Three single launches of every (variant, k in {4, kmax}) on the headline
    input, in the order printed; rocprofv3 --pmc over this run gives each
    configuration's FETCH_SIZE / WRITE_SIZE per dispatch."""
    import torch
    from hindemith_amd import kernels as K
    from hindemith_amd import synthetic as S
    taps = S.hs_stencil_taps()
    I1h, I2h, _, _ = S.hs_pair(H, W, seed=0, flow="affine")
    I1, I2 = torch.from_numpy(I1h).cuda(), torch.from_numpy(I2h).cuda()
    ws = torch.empty((5, H, W), device="cuda")
    u = torch.zeros((H, W), device="cuda")
    v = torch.zeros_like(u)
    Ix, Iy, It = K.hs_setup(I1, I2, taps["Dx"], taps["Dy"], taps["Dt"], out=ws[0:3])
    order = []
    for vi, info in enumerate(K.hs_variants()):
        for k in sorted({min(4, info["kmax"]), info["kmax"]}):
            for _ in range(3):
                K.hs_jacobi(Ix, Iy, It, 1.0, taps["Avg"], k, u, v, work=ws[3:5], variant=vi,
                            k=k)
            torch.cuda.synchronize()
            order.append({"variant": vi, "k": k, "kernel": info["kernel"], "launches": 3})
    print(json.dumps({"H": H, "W": W, "order": order}))


def cpu_baseline(I1, I2, taps, H, W):
    import oracle
    oracle.build_oracle()
    threads = oracle.num_threads()
    # the timed step's own workload (one solve: setup + all iterations), about
    # 10 s of CPU work in all
    it, runs = 500, 7
    oracle.hs_solve(I1, I2, 1.0, taps, 10)  # warm-up
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.hs_solve(I1, I2, 1.0, taps, it)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": round(H * W * it / med / 1e6, 2), "unit": "Mpixel*iter/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle C/OpenMP restatement, {W}x{H} setup + {it} Jacobi iterations "
                      f"(one full solve), median of {runs} runs after 1 warm-up "
                      f"({med:.2f} s each, {sum(times):.1f} s in all)"}


def endpoint_error(u, v, fx, fy, margin=8):
    """Mean endpoint error of the flow against the true synthetic flow, over the
    interior (margin px from each border)."""
    sl = (slice(margin, -margin), slice(margin, -margin))
    du = np.asarray(u, np.float64)[sl] - np.asarray(fx, np.float64)[sl]
    dv = np.asarray(v, np.float64)[sl] - np.asarray(fy, np.float64)[sl]
    return float(np.mean(np.sqrt(du * du + dv * dv)))


def _time_ms(fn, reps, warm=2):
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _time_graph_ms(fn, launches=50, reps=5):
    """Per-call GPU time of fn: launches calls captured in one hipGraph, the
    graph replayed reps times, median of replay time / launches.  Removes the
    host-side cost of each call, which exceeds the kernel time of a small
    bandwidth kernel."""
    import torch
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            for _ in range(launches):
                fn()
    torch.cuda.current_stream().wait_stream(s)
    g.replay()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / launches)
    del g
    return float(np.median(times))


def extra_4k(K, S, taps):
    """3840x2160 fp32, 500 iterations, affine flow (same kernel path)."""
    import torch
    H, W, iters = 2160, 3840, 500
    I1h, I2h, _, _ = S.hs_pair(H, W, seed=1, flow="affine")
    I1, I2 = torch.from_numpy(I1h).cuda(), torch.from_numpy(I2h).cuda()
    ws = torch.empty((5, H, W), device="cuda")
    u = torch.zeros((H, W), device="cuda")
    v = torch.zeros_like(u)
    Ix, Iy, It = K.hs_setup(I1, I2, taps["Dx"], taps["Dy"], taps["Dt"], out=ws[0:3])
    vi, k = K.hs_tune(Ix, Iy, It, 1.0, taps["Avg"])

    def solve():
        u.zero_()
        v.zero_()
        K.hs_solve(I1, I2, 1.0, taps, iters, u, v, workspace=ws)
    ms = _time_ms(solve, 3, warm=1)
    val = H * W * iters / (ms * 1e-3) / 1e6
    return {"workload": f"{W}x{H} fp32 {iters} it", "value": round(val, 1),
            "unit": "Mpixel*iter/s", "ms_per_solve": round(ms, 3), "variant": vi, "k": k,
            "hbm_roofline_fraction": round(val / (HBM_PEAK_GBS * 1e3 / BYTES_PER_PX_ITER), 4)}


def extra_bandwidth(K, S):
    """2048x2048 bandwidth kernels, 200 repeats each (GB/s of algorithmic
    bytes), timed as hipGraph replays of 50 launches."""
    import torch
    H = W = 2048
    N = H * W
    rng = np.random.default_rng(0)
    x = torch.from_numpy(rng.uniform(-1, 1, (H, W)).astype(np.float32)).cuda()
    y = torch.empty_like(x)
    out = {}
    lap = S.laplacian5_taps()
    ms = _time_graph_ms(lambda: K.stencil2d(x, lap, out=y))
    out["stencil_5pt"] = {"us": round(ms * 1e3, 2), "GBps": round(8 * N / ms / 1e6, 1)}
    planes, ox, oy = S.laplacian5_dia_planes(H, W)
    A = torch.from_numpy(planes).cuda()
    ms = _time_graph_ms(lambda: K.dia2d_spmv(A, ox, oy, x, out=y))
    out["dia2d_5diag"] = {"us": round(ms * 1e3, 2), "GBps": round(28 * N / ms / 1e6, 1)}
    # the same operator as a 1-D scipy-style dia_matrix of size N (offsets 0, +-1, +-W)
    off1 = np.array([0, -1, 1, -W, W], np.int32)
    D1 = torch.from_numpy(np.ascontiguousarray(planes.reshape(5, N))).cuda()
    ms = _time_graph_ms(lambda: K.dia_spmv(D1, off1, x.reshape(-1), out=y.reshape(-1)))
    out["dia1d_5diag"] = {"us": round(ms * 1e3, 2), "GBps": round(28 * N / ms / 1e6, 1)}
    del D1
    # sum2d: deterministic two-pass reduction, 4 B per element read
    part = torch.empty(1024, device="cuda", dtype=torch.float64)
    s1 = torch.empty(1, device="cuda")
    ms = _time_graph_ms(lambda: K.sum2d(x, partials=part, out=s1))
    out["sum2d"] = {"us": round(ms * 1e3, 2), "GBps": round(4 * N / ms / 1e6, 1)}
    ip, ix, d = S.dia2d_to_csr(planes, ox, oy)
    tip, tix, td = (torch.from_numpy(a).cuda() for a in (ip, ix, d))
    xf = x.reshape(-1)
    yf = y.reshape(-1)
    nnz = int(ip[-1])
    ms = _time_graph_ms(lambda: K.csr_spmv(tip, tix, td, xf, out=yf))
    out["csr_5pt"] = {"nnz": nnz, "us": round(ms * 1e3, 2),
                      "GBps": round((8 * nnz + 12 * N) / ms / 1e6, 1)}
    del tip, tix, td
    ip, ix, d = S.powerlaw_csr(N, np.random.default_rng(1))
    tip, tix, td = (torch.from_numpy(a).cuda() for a in (ip, ix, d))
    nnz = int(ip[-1])
    rb = K.CSRRowBlocks(ip, "cuda")  # built once per matrix, on the host
    ms = _time_graph_ms(lambda: K.csr_spmv_rowblocks(tip, tix, td, xf, rb, out=yf))
    bins = K.CSRBins(ip, "cuda")
    ms_bin = _time_graph_ms(lambda: K.csr_spmv_binned(tip, tix, td, xf, bins, out=yf),
                            launches=20)
    ms_vec = _time_graph_ms(lambda: K.csr_spmv(tip, tix, td, xf, out=yf), launches=20)
    # column-sliced form: the slicing is built once per sparsity pattern (host)
    # and the values permuted once per matrix (device), both outside the time
    sl = K.CSRSlices(ip, ix, N, "cuda")
    sd = sl.values(td)
    ms_sl = _time_graph_ms(lambda: K.csr_spmv_sliced(sl, sd, xf, out=yf))
    best = min((ms, "row blocks (<= 2048 entries per work-group)"),
               (ms_sl, "column slices (one 2.1 MB range of x per XCD) + slice-sum pass"))
    out["csr_powerlaw_zipf1.5"] = {"nnz": nnz, "us": round(best[0] * 1e3, 2),
                                   "GBps": round((8 * nnz + 12 * N) / best[0] / 1e6, 1),
                                   "form": best[1],
                                   "us_row_blocks": round(ms * 1e3, 2),
                                   "us_column_sliced": round(ms_sl * 1e3, 2),
                                   "us_row_binned": round(ms_bin * 1e3, 2),
                                   "us_unbinned": round(ms_vec * 1e3, 2)}
    del sl, sd
    return out


def extra_batched(S):
    """64 uneven pairs (sides U(64,1024)), 200 iterations: one persistent
    graph-captured launch vs 64 plain per-pair solves.  Plain baselines: the
    temporally blocked solve of each pair called eagerly (speedup) and the same
    calls captured in one hipGraph (speedup_vs_plain_graph, GPU time only), and
    each pair with one fused Jacobi launch per iteration, the reference's launch
    structure (speedup_vs_launch_per_iteration)."""
    import torch
    from hindemith_amd.batched import BatchedHS
    pairs = S.batched_pairs(64, seed=0)
    B = BatchedHS([torch.from_numpy(p["I1"]).cuda() for p in pairs],
                  [torch.from_numpy(p["I2"]).cuda() for p in pairs],
                  [p["alpha"] for p in pairs], iters=200)
    vi, k = B.tune()
    B.run(graph=True, sync=True)
    # both sides timed over ~0.1-0.2 s after warm-up calls, at the sustained rate
    ms_b = _time_ms(lambda: B.run(graph=True), 20, warm=5)
    ms_p = _time_ms(B.solve_each, 6, warm=2)
    # the same 64 solves captured in one hipGraph: GPU time without host cost
    ms_pg = _time_graph_ms(B.solve_each, launches=1, reps=3)
    ms_1 = _time_ms(lambda: B.solve_each(variant=0, k=1), 2, warm=1)
    px = sum(p["I1"].size for p in pairs)
    return {"pairs": 64, "megapixels": round(px / 1e6, 2), "iters": 200,
            "batched_pairs_per_s": round(64 / (ms_b * 1e-3), 1),
            "plain_pairs_per_s": round(64 / (ms_p * 1e-3), 1),
            "speedup": round(ms_p / ms_b, 2), "batched_ms": round(ms_b, 3),
            "plain_ms": round(ms_p, 3), "plain_graph_ms": round(ms_pg, 3),
            "speedup_vs_plain_graph": round(ms_pg / ms_b, 2),
            "plain_launch_per_iteration_ms": round(ms_1, 3),
            "speedup_vs_launch_per_iteration": round(ms_1 / ms_b, 2),
            "batched_mpx_iter_per_s": round(px * 200 / (ms_b * 1e-3) / 1e6, 1),
            "variant": vi, "k": k}


def extra_hscg(K, S):
    """The preconditioned-CG operations, the Brox SOR blocks and the image warp at
    1920x1080 fp32, timed as hipGraph replays: the warp of frame 2 by the synthetic
    flow (16 B/px:
    read src, offx, offy; write dst), the 2x2 block-Jacobi preconditioner (28 B/px: read diag0, diag1,
    offdiag, x0, x1; write y0, y1) and 8 HS operator powers (algorithmic bytes
    as the reference's one launch per power: V[0] copy 16 B/px + 24 B/px per
    power = 184 B/px; the blocked kernel's minimum traffic is 88 B/px: read Ix,
    Iy, r once, write the 8 planes)."""
    import torch
    H, W, steps = 1080, 1920, 8
    n = H * W
    taps = S.hs_stencil_taps()
    I1, I2, _, _ = S.hs_pair(H, W, seed=0, flow="affine")
    Ix, Iy, It = K.hs_setup(torch.from_numpy(I1).cuda(), torch.from_numpy(I2).cuda(),
                            taps["Dx"], taps["Dy"], taps["Dt"])
    r0, r1 = -(Ix * It), -(Iy * It)
    d0, d1, od = Ix * Ix, Iy * Iy, Ix * Iy
    y = (torch.empty_like(r0), torch.empty_like(r0))
    lap = S.laplacian5_taps()
    ms_p = _time_graph_ms(lambda: K.hs_precondition(lap, d0, d1, od, r0, r1, out=y))
    V = (torch.empty((steps, H, W), device="cuda"), torch.empty((steps, H, W), device="cuda"))
    ms_m = _time_graph_ms(lambda: K.hs_matrix_powers(1.0, Ix, Iy, r0, r1, steps, out=V),
                          launches=10)
    alg = (16 + 24 * (steps - 1)) * n
    fx, fy = (torch.as_tensor(f, device="cuda", dtype=torch.float32).contiguous()
              for f in S.affine_flow(H, W))
    wdst = torch.empty_like(r0)
    I2d = torch.from_numpy(I2).cuda()
    ms_w = _time_graph_ms(lambda: K.warp_img2d(I2d, fx, fy, out=wdst))
    # Brox SOR blocks: the psi-weighted Laplacian into 5 planes (24 B/px: read psi,
    # write 5 planes) and the 2x2 preconditioner on its diagonal (32 B/px)
    Ab = torch.zeros((5, H, W), device="cuda")
    psi = (Ix * Ix + 1.0).contiguous()
    ms_bs = _time_graph_ms(lambda: K.brox_set_matrix(Ab, psi, 0.5))
    yb = (torch.empty_like(r0), torch.empty_like(r0))
    ms_bp = _time_graph_ms(lambda: K.brox_precondition(Ab, d0, d1, od, r0, r1, out=yb))
    del Ab
    return {"workload": f"{W}x{H} fp32",
            "warp_img2d": {"us": round(ms_w * 1e3, 2), "GBps": round(16 * n / ms_w / 1e6, 1)},
            "brox_set_matrix": {"us": round(ms_bs * 1e3, 2),
                                "GBps": round(24 * n / ms_bs / 1e6, 1)},
            "brox_precondition": {"us": round(ms_bp * 1e3, 2),
                                  "GBps": round(32 * n / ms_bp / 1e6, 1)},
            "precondition": {"us": round(ms_p * 1e3, 2), "GBps": round(28 * n / ms_p / 1e6, 1)},
            "matrix_powers_8": {"us": round(ms_m * 1e3, 2),
                                "Mpx_power_per_s": round(n * (steps - 1) / ms_m / 1e3, 1),
                                "GBps_algorithmic": round(alg / ms_m / 1e6, 1),
                                "GBps_min_traffic": round(88 * n / ms_m / 1e6, 1)}}


def extra_frontend(S):
    """The drop-in @hm path: the canonical Horn-Schunck program at 1920x1080,
    500 iterations, called with host numpy arrays, so the PCIe copies of the
    frames and of the returned flow are inside the time (wall clock per call,
    median of 3 after a warm-up call that compiles the record)."""
    os.environ.setdefault("HM_GPU0", "1")
    from hindemith_amd import hindemith as Hm
    Hm.defines.update(nd20=1, nd21=1, nt20=8, nt21=8, ne20=2, ne21=2)
    hm, Stencil = Hm.hm, Hm.Stencil

    @hm
    def hs_1080p(I1, I2, Dx, Dy, Dt, Avg, alpha2, u, v):
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

    taps = S.hs_stencil_taps()
    H, W = 1080, 1920
    I1, I2, _, _ = S.hs_pair(H, W, seed=0, flow="affine")
    st = {k: Stencil(*taps[k]) for k in ("Dx", "Dy", "Dt", "Avg")}

    def call():
        z = np.zeros_like(I1)
        return hs_1080p(I1=I1, I2=I2, alpha2=np.float32(1.0), u=z, v=z.copy(), **st)

    call()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        call()
        times.append(time.perf_counter() - t0)
    sec = float(np.median(times))
    return {"workload": f"@hm canonical HS program, {W}x{H} fp32 500 it, host arrays",
            "value": round(H * W * 500 / sec / 1e6, 1), "unit": "Mpixel*iter/s",
            "ms_per_call": round(sec * 1e3, 3),
            "plan": [l for l in hs_1080p.plan().splitlines() if l.strip()][:4]}


def extra_adjoint(S):
    """8 pairs at 512x512, 200 iterations, checkpoints every 20: time of the
    hand-written backward vs the forward."""
    import torch
    from hindemith_amd.autograd import horn_schunck
    rng = np.random.default_rng(0)
    I1, I2, fx, fy = [], [], [], []
    for b in range(8):
        a, c, x, y = S.hs_pair(512, 512, seed=100 + b, flow="translation",
                               trans=tuple(rng.uniform(-3, 3, 2)))
        I1.append(a), I2.append(c), fx.append(x), fy.append(y)
    tI1 = torch.tensor(np.stack(I1), device="cuda", requires_grad=True)
    tI2 = torch.tensor(np.stack(I2), device="cuda", requires_grad=True)
    al = torch.tensor(rng.uniform(0.5, 2, 8).astype(np.float32), device="cuda",
                      requires_grad=True)
    tfx = torch.tensor(np.stack(fx), device="cuda")
    tfy = torch.tensor(np.stack(fy), device="cuda")
    st = {}

    def fwd():
        st["u"], st["v"] = horn_schunck(tI1, tI2, al, iters=200, ckpt_every=20)

    def bwd():
        loss = ((st["u"] - tfx) ** 2 + (st["v"] - tfy) ** 2).mean()
        loss.backward()

    def loop(n, with_bwd):
        for _ in range(n):
            fwd()
            if with_bwd:
                bwd()

    # Each figure comes from passes enqueued back to back (nothing on the path
    # waits for the host), so neither side starts from an idle GPU: the
    # forward alone, then forward + backward pairs; the backward is the
    # difference.  Repeated, median.
    loop(3, True)
    torch.cuda.synchronize()
    n, t_f, t_fb = 5, [], []
    for _ in range(3):
        for with_bwd, acc in ((False, t_f), (True, t_fb)):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            loop(n, with_bwd)
            e1.record()
            torch.cuda.synchronize()
            acc.append(e0.elapsed_time(e1) / n)
    f = float(np.median(t_f))
    b = float(np.median(t_fb)) - f
    return {"workload": "8 x 512x512, 200 it, ckpt every 20", "forward_ms": round(f, 3),
            "backward_ms": round(b, 3), "backward_over_forward": round(b / f, 2)}


def load_counters(variant, k, H, W, iters):
    """Per-launch rocprofv3 counters of the fused Jacobi kernel from the
    committed summary, if they were measured for this exact solve (variant, k,
    H, W, iters): the mean over every launch of a solve."""
    try:
        d = json.load(open(COUNTERS))
    except (OSError, ValueError):
        return None
    if (d.get("H"), d.get("W"), d.get("variant"), d.get("k"), d.get("iters")) != \
            (H, W, variant, k, iters):
        return None
    return d


def hs_roofline(vinfo, variant, k, H, W, iters, launches, jac_ms):
    """Roofline of the fused Jacobi kernel over one solve's launches.

    The binding resource is the vector ALU: the blocked kernel moves about a
    seventh of its algorithmic bytes through HBM, and its launches are spent
    issuing VALU instructions (the step loop) and waiting on the load phase.
    achieved = issued VALU instructions per launch (SQ_INSTS_VALU, counted per
    wave) x 64 lanes / the mean launch time measured live with HIP events;
    peak = the non-packed issue rate VALU_LANE_PEAK.  Also reported: the same
    launches against HBM (algorithmic bytes, the figure the 28 B/px.iter
    definition gives, and the measured FETCH/WRITE traffic) and against LDS."""
    t = jac_ms * 1e-3 / launches
    alg = BYTES_PER_PX_ITER * H * W * iters // launches
    hbm = {"algorithmic_bytes_per_launch": alg,
           "achieved_algorithmic_GBps": round(alg / t / 1e9, 1),
           "frac_algorithmic": round(alg / t / 1e9 / HBM_PEAK_GBS, 4)}
    base = {"kernel": vinfo["kernel"], "launches_per_step": launches,
            "avg_launch_us": round(t * 1e6, 2)}
    d = load_counters(variant, k, H, W, iters)
    if d is None:  # counters not measured for this solve: the HBM roof only
        return dict({"bound": "hbm", "achieved": hbm["achieved_algorithmic_GBps"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm["frac_algorithmic"],
                     "traffic": None, "hbm": hbm}, **base)
    c = d["per_launch"]
    traffic = int(c["FETCH_SIZE"] * 1024 * d["fetch_scale"] + c["WRITE_SIZE"] * 1024 *
                  d["write_scale"])
    hbm.update({"traffic_bytes_per_launch": traffic,
                "achieved_measured_GBps": round(traffic / t / 1e9, 1),
                "frac_measured": round(traffic / t / 1e9 / HBM_PEAK_GBS, 4)})
    lanes = c["SQ_INSTS_VALU"] * 64
    lds_bytes = (c["SQ_INSTS_LDS_LOAD_BANDWIDTH"] + c["SQ_INSTS_LDS_STORE_BANDWIDTH"]) * 64
    lds = {"bytes_per_launch": int(lds_bytes), "achieved_TBps": round(lds_bytes / t / 1e12, 2),
           "peak_TBps": round(LDS_PEAK_BPS / 1e12, 1),
           "frac": round(lds_bytes / t / LDS_PEAK_BPS, 4)}
    return dict({"bound": "valu", "achieved": round(lanes / t / 1e12, 2),
                 "peak": round(VALU_LANE_PEAK / 1e12, 2), "unit": "Tlane-instr/s",
                 "frac": round(lanes / t / VALU_LANE_PEAK, 4), "traffic": traffic,
                 "valu_insts_per_launch": int(c["SQ_INSTS_VALU"]),
                 "fp32_tflops": round(c["SQ_INSTS_VALU_FLOPS_FP32"] * 64 / t / 1e12, 2),
                 "counters": os.path.relpath(COUNTERS, HERE), "hbm": hbm, "lds": lds}, **base)


def main():
    args = parse()
    if args.copy_mb:
        copy_calibration(args.copy_mb)
        return 0
    if args.traffic_sweep:
        traffic_sweep(args.height, args.width)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (the "
                                   "multi-GPU design is domain decomposition, not built here)"}))
        return 2
    import torch
    from hindemith_amd import kernels as K
    from hindemith_amd import synthetic as S

    H, W, iters = args.height, args.width, args.iters
    taps = S.hs_stencil_taps()
    if args.secondary:
        fn = {"4k": lambda: extra_4k(K, S, taps), "bandwidth": lambda: extra_bandwidth(K, S),
              "batched": lambda: extra_batched(S), "adjoint": lambda: extra_adjoint(S),
              "frontend": lambda: extra_frontend(S), "hscg": lambda: extra_hscg(K, S)}
        print(json.dumps({args.secondary: fn[args.secondary]()}))
        return 0
    I1h, I2h, fx, fy = S.hs_pair(H, W, seed=0, flow="affine")
    dev = torch.device("cuda:0")
    I1 = torch.from_numpy(I1h).to(dev)
    I2 = torch.from_numpy(I2h).to(dev)
    uv = torch.zeros((2, H, W), device=dev)  # u and v, cleared by one fill per step
    u, v = uv[0], uv[1]
    ws = torch.empty((5, H, W), device=dev)
    Ix, Iy, It = ws[0], ws[1], ws[2]
    work = ws[3:5]

    K.hs_setup(I1, I2, taps["Dx"], taps["Dy"], taps["Dt"], out=ws[0:3])
    if args.variant is not None:
        variant, k = args.variant, args.k or K.hs_variants()[args.variant]["kmax"]
    elif args.no_tune:
        variant, k = K.hs_config(H, W)
    else:
        variant, k = K.hs_tune(Ix, Iy, It, 1.0, taps["Avg"])

    def step(ev=None):
        uv.zero_()
        K.hs_setup(I1, I2, taps["Dx"], taps["Dy"], taps["Dt"], out=ws[0:3])
        if ev is not None:
            ev[0].record()
        K.hs_jacobi(Ix, Iy, It, 1.0, taps["Avg"], iters, u, v, work=work, variant=variant, k=k)
        if ev is not None:
            ev[1].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    graph_ok = None
    if not args.no_graph:
        # One step (zero u, v; setup; every Jacobi launch) captured once and
        # replayed; the replayed flow must equal the eager flow bit for bit.
        step()
        torch.cuda.synchronize()
        ref_u, ref_v = u.clone(), v.clone()
        cap_ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                step(cap_ev)
        torch.cuda.current_stream().wait_stream(s)
        graph_ok = True
        for _ in range(3):
            g.replay()
            junk = torch.randn(4096, device=dev).sum()  # eager work between replays
            torch.cuda.synchronize()
            graph_ok = graph_ok and bool(torch.equal(u, ref_u) and torch.equal(v, ref_v))
            del junk
        run_step = lambda i: (evs[i][0].record(), g.replay(), evs[i][1].record())  # noqa: E731
    else:
        run_step = lambda i: step(evs[i])  # noqa: E731
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        run_step(i)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    jac_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    if not args.no_graph:
        # the replayed events bracket the whole step; subtract the setup kernel's
        # share measured eagerly so the per-launch figure is the Jacobi kernel's
        setup_ms = _time_ms(lambda: K.hs_setup(I1, I2, taps["Dx"], taps["Dy"], taps["Dt"],
                                               out=ws[0:3]), 5)
        jac_ms = max(jac_ms - setup_ms, 1e-6)

    vinfo = K.hs_variants()[variant]
    launches = -(-iters // k)
    ms_per_step = elapsed * 1e3 / args.steps
    value = H * W * iters / (ms_per_step * 1e-3) / 1e6
    # the mean launch carries iters / launches iterations (a solve of 500 at
    # k = 8 is 62 launches of 8 and one of 4)
    roofline = hs_roofline(vinfo, variant, k, H, W, iters, launches, jac_ms)

    u_h = u.cpu().numpy()
    v_h = v.cpu().numpy()
    epe = endpoint_error(u_h, v_h, fx, fy)

    out = {
        "metric": "Mpixel*iter/s (HS, 1 GPU)", "value": round(value, 1),
        "unit": "Mpixel*iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: 32 Gaussian blobs + N(0,0.01) noise, frame 2 = affine warp "
                "(rot 0.5 deg, scale 1.01, shift (2,-1) px)",
        "config": {"workload": f"horn_schunck_jacobi_{W}x{H}_fp32_{iters}it", "height": H,
                   "width": W, "iters": iters, "alpha": 1.0,
                   "kernel_variant": variant, "iters_per_launch": k,
                   "tile": [vinfo["tile_w"], vinfo["tile_h"]],
                   "hbm_roofline_fraction_of_metric": round(value / (HBM_PEAK_GBS * 1e9 /
                                                                     BYTES_PER_PX_ITER / 1e6), 4),
                   "endpoint_error_px": round(epe, 4),
                   "hip_graph": not args.no_graph, "graph_replay_bitexact": graph_ok},
        "roofline": roofline,
    }
    if not args.no_extras:
        sec = {}
        for name, fn in (("hs_3840x2160", lambda: extra_4k(K, S, taps)),
                         ("bandwidth_2048x2048", lambda: extra_bandwidth(K, S)),
                         ("batched_vs_plain", lambda: extra_batched(S)),
                         ("adjoint_vs_forward", lambda: extra_adjoint(S)),
                         ("hm_frontend_host_arrays", lambda: extra_frontend(S)),
                         ("hs_cg_operations", lambda: extra_hscg(K, S))):
            try:
                sec[name] = fn()
            except Exception as e:  # a secondary workload must not hide the headline
                sec[name] = {"error": f"{type(e).__name__}: {e}"}
            print(f"[bench] {name}: {json.dumps(sec[name])}", file=sys.stderr, flush=True)
        out["secondary"] = sec
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(I1h, I2h, taps, H, W)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
