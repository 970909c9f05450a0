#!/usr/bin/env python3
"""Benchmark of the adjoint NFFT on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-bootstrap]

Workload (the headline configuration): plain adjoint NFFT, FP32 library,
accurate mode (sigma = 2, m = 8), N = 1e6 clustered points (80 % in 50
Gaussian clumps of width 0.005, 20 % uniform), y = 3 sinusoids + noise,
M = 2^20 modes (grid G = 2^21).  A step is one transform of one y with the
plan (points binned once per x, as the reference precomputes its filter per x)
reused; inputs are resident in HBM.  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "adj-NFFT/s at N=1e6,M=2^20 FP32 + bootstrap LCs/s vs plain, at rel err ≤1e-5"
HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E


class HipEvents:
    """hipEvent_t handles through libamdhip64 (timing on the launch stream)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]

    def create(self, n):
        out = []
        for _ in range(n):
            e = ctypes.c_void_p()
            assert self.h.hipEventCreate(ctypes.byref(e)) == 0
            out.append(e.value)
        return out

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def ms(self, a, b):
        v = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(v), a, b) == 0
        return v.value

    def destroy(self, evs):
        for e in evs:
            self.h.hipEventDestroy(e)


def read_traffic(kernel: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), scaled by a copy kernel of known size; kept in profiles/."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_c2_fp32.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch") and kernel in line:
            try:
                return float(line.split()[-1])
            except ValueError:
                return None
    return None


def cpu_baseline(x, y, M, seconds):
    """The oracle's FP64 Greengard-Lee spread + FFT (same method), OpenMP on the
    host cores, timed on repetitions of the same C2 input for about `seconds`;
    plus the direct NDFT on a 4096-frequency subset, scaled to M modes."""
    import oracle

    xd = x.astype(np.float64)
    yd = y.astype(np.float64)
    oracle.gl_adjoint(xd[:1000], yd[:1000], 64)  # load + warm
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.gl_adjoint(xd, yd, M, 2, 8)
        reps += 1
        el = time.perf_counter() - t0
        if el > seconds or reps >= 50:
            break
    rate = reps / el
    k0 = np.arange(-2048, 2048, 1024, dtype=np.int64)
    t1 = time.perf_counter()
    oracle.ndft_bands(xd, yd, k0, 1024)
    tn = time.perf_counter() - t1
    ndft_rate = 1.0 / (tn * M / 4096.0)
    return {
        "value": rate,
        "unit": "adj-NFFT/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"{reps} transforms of the same workload (N=1e6 clustered, M=2^20) by the "
                  f"CPU oracle: FP64 Gaussian spread + own radix-2 FFT, OpenMP, {el:.1f} s",
        "ndft_value": ndft_rate,
        "ndft_sample": f"direct FP64 NDFT of N=1e6 points on 4096 frequencies ({tn:.2f} s), "
                       f"scaled to M=2^20",
        "cpu_model": _cpu_model(),
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_bandwidth(torch):
    n = 1 << 28  # 1 GiB of float32
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    s = e0.elapsed_time(e1) / 1e3 / 10
    del a, b
    return 2 * 4 * n / s


def multi_rhs(plan, yd, out, M, N, info, nrhs, stream, ev, outer, K):
    """nrhs value vectors of the same plan per call (cuna_nfft_adjoint_many)."""
    import torch

    sp = stream.cuda_stream
    ym = torch.stack([yd * (1.0 + 0.125 * r) for r in range(nrhs)]).contiguous()
    outm = torch.empty((nrhs, M), dtype=torch.complex64, device="cuda")
    plan.adjoint_many(ym, out=outm, stream=stream)
    stream.synchronize()
    Km = max(4, K // nrhs)
    ev.record(outer[0], sp)
    for _ in range(Km):
        plan.adjoint_many(ym, out=outm, stream=stream)
    ev.record(outer[1], sp)
    stream.synchronize()
    ms = ev.ms(outer[0], outer[1]) / Km
    diff = float(torch.linalg.vector_norm(outm[0] - out[0]) / torch.linalg.vector_norm(out[0]))
    return {
        "what": f"{nrhs} value vectors of the same plan per call (cuna_nfft_adjoint_many): one "
                "spread launch and one batched FFT; not the headline, which is one vector per call",
        "nrhs": nrhs, "ms_per_call": ms, "adj_nfft_per_s": nrhs * 1e3 / ms,
        "first_row_rel_diff_vs_single": diff,
        "frac_of_hbm_peak": nrhs * (N * (4 + 4) + 2 * info.G * 4 + M * 8) / (ms * 1e-3) / HBM_PEAK,
    }


def bench_c2(args, torch, cu):
    from tests._data import clustered_points, rel_l2
    import oracle

    N, M = 1_000_000, 1 << 20
    x, y = clustered_points(N, 2024)
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(y).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = cu.NFFTPlan(xd, M)
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    info = plan.info()
    out = torch.empty((1, M), dtype=torch.complex64, device="cuda")
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    for _ in range(args.warmup):
        plan.adjoint(yd, out, stream=stream)
    stream.synchronize()
    # plan + execute with the code objects already loaded: a new plan over
    # the same points, then one transform (plan_ms above includes the first
    # launches' module loading)
    warm = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p2 = cu.NFFTPlan(xd, M)
        t1 = time.perf_counter()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        p2.adjoint(yd, out, stream=stream)
        stream.synchronize()
        warm.append(((t1 - t0) * 1e3, (t2 - t0) * 1e3, (time.perf_counter() - t0) * 1e3))
        p2.close()
    warm.sort(key=lambda w: w[2])
    plan_warm = {"host_return_ms": warm[1][0], "plan_ms": warm[1][1],
                 "plan_and_execute_ms": warm[1][2]}
    ev = HipEvents()
    K = args.steps
    L = plan.L
    outer = ev.create(2)
    # host enqueue cost of one step (ctypes call + launches)
    t0 = time.perf_counter()
    for _ in range(10):
        plan.adjoint(yd, out, stream=stream)
    host_step = (time.perf_counter() - t0) / 10
    stream.synchronize()
    # a spin kernel holds the stream while the host enqueues all K steps, so
    # the GPU then runs them back to back; ms per spin cycle calibrated first
    with torch.cuda.stream(stream):
        ev.record(outer[0], sp)
        torch.cuda._sleep(10_000_000)
        ev.record(outer[1], sp)
    stream.synchronize()
    ms_per_cycle = ev.ms(outer[0], outer[1]) / 1e7
    hold_ms = 1.5 * host_step * 1e3 * (K + 8) + 20.0

    def held(fn):
        """Enqueue K steps behind a spin kernel; returns (ms of the K steps,
        host enqueue ms)."""
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            torch.cuda._sleep(int(hold_ms / ms_per_cycle))
        ev.record(outer[0], sp)
        th = time.perf_counter()
        for k in range(K):
            fn(k)
        enq = (time.perf_counter() - th) * 1e3
        ev.record(outer[1], sp)
        stream.synchronize()
        torch.cuda.synchronize()
        return ev.ms(outer[0], outer[1]), enq

    # the timed region: K transforms back to back, nothing else on the stream
    total_ms, enqueue_ms = held(lambda k: plan.adjoint(yd, out, stream=stream))
    ms_step = total_ms / K
    host_bound = enqueue_ms > hold_ms
    # a second pass with profiling events (per-kernel durations for the
    # roofline): the spread's start and stop are taken by its own dispatch;
    # the FFT passes are bracketed by event markers, which add small gaps, so
    # this pass is not `value`
    stage = ev.create(4 * K)

    def staged(k):
        arr = (ctypes.c_void_p * 4)(*stage[4 * k:4 * k + 4])
        L.cuna_nfft_set_profile_events(plan.h, arr, 4)
        plan.adjoint(yd, out, stream=stream)

    held(staged)
    L.cuna_nfft_set_profile_events(plan.h, None, 0)
    spread = np.mean([ev.ms(stage[4 * k], stage[4 * k + 1]) for k in range(K)])
    colp = np.mean([ev.ms(stage[4 * k + 1], stage[4 * k + 2]) for k in range(K)])
    rowp = np.mean([ev.ms(stage[4 * k + 2], stage[4 * k + 3]) for k in range(K)])
    ev.destroy(stage)
    # the same K transforms replayed from a captured graph (no events)
    graph = cu.capture(lambda: [plan.adjoint(yd, out, stream=stream) for _ in range(K)], stream)
    graph.launch(stream)
    stream.synchronize()
    ev.record(outer[0], sp)
    graph.launch(stream)
    ev.record(outer[1], sp)
    stream.synchronize()
    graph_ms = ev.ms(outer[0], outer[1]) / K
    graph.close()
    # several right-hand sides of the same plan per call (cuna_nfft_adjoint_many):
    # one spread launch and one batched FFT for all of them
    multi = None
    nrhs = 8 if not args.no_multi_rhs else 0
    if nrhs:
        multi = multi_rhs(plan, yd, out, M, N, info, nrhs, stream, ev, outer, K)
    # eager launches from Python, for reference (host-bound at this size)
    ev.record(outer[0], sp)
    for _ in range(K):
        plan.adjoint(yd, out, stream=stream)
    ev.record(outer[1], sp)
    stream.synchronize()
    eager_ms = ev.ms(outer[0], outer[1]) / K
    ev.destroy(outer)

    # accuracy of this output vs the FP64 direct NDFT on frequency bands
    got = out[0].cpu().numpy()
    rng = np.random.default_rng(1)
    k0 = np.concatenate([[-M // 2, -16], rng.integers(-M // 2, M // 2 - 32, 6)]).astype(np.int64)
    ref = oracle.ndft_bands(x.astype(np.float64), y.astype(np.float64), k0, 32)
    sel = np.stack([got[k + M // 2:k + M // 2 + 32] for k in k0])
    err = rel_l2(sel, ref)

    G = info.G
    bytes_total = N * (4 + 4) + 2 * G * 4 + M * 8
    kernels = {
        "spread (k_spread_wv)": (spread, N * (4 + 4) + G * 4, "k_spread_wv"),
        "fft pass 1 (k_fft_col)": (colp, G * 4 + (G // 2) * 8, "k_fft_col"),
        "fft pass 2 + deconvolution (k_fft_row_real)": (rowp, (G // 2) * 8 + M * 8,
                                                         "k_fft_row_real"),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    t_dom, b_dom, kname = kernels[dom]
    achieved = b_dom / (t_dom * 1e-3)
    return {
        "ms_per_step": ms_step,
        "value": 1e3 / ms_step,
        "eager_ms_per_step": eager_ms,
        "graph_ms_per_step": graph_ms,
        "timing_host_bound": host_bound,
        "plan_ms": plan_ms,
        "plan_warm": plan_warm,
        "rel_err_vs_ndft": err,
        "stages_ms": {k: v[0] for k, v in kernels.items()},
        "multi_rhs": multi,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": read_traffic(kname),
            "algorithmic_bytes_per_launch": b_dom,
            "avg_launch_ms": t_dom,
        },
        "pipeline": {
            "algorithmic_bytes_per_transform": bytes_total,
            "achieved_GBps": bytes_total / (ms_step * 1e-3) / 1e9,
            "frac_of_hbm_peak": bytes_total / (ms_step * 1e-3) / HBM_PEAK,
        },
        "info": {"G": G, "tile_log2": info.tile_log2, "ntiles": info.ntiles, "m": info.m},
        "x": x,
        "y": y,
        "M": M,
    }


def hbm_roofline(nbytes, ms, copy_bw):
    """Compulsory bytes over a measured time, against the HBM peak and the
    box's own measured device-to-device copy bandwidth."""
    gbps = nbytes / (ms * 1e-3)
    return {"bound": "hbm", "compulsory_bytes": nbytes, "ms": ms, "achieved_GBps": gbps / 1e9,
            "frac_of_hbm_peak": gbps / HBM_PEAK, "frac_of_copy_bw": gbps / copy_bw}


def bench_bootstrap(args, torch, cu, copy_bw):
    """Batched bootstrap (the batching configuration): 1024 light curves of
    log-uniform length in [100, 5e4] (8.2e6 points), times in days, FP64 phase
    reduction with df = 1/3650 c/d, M = 2^14, 100 replicates per curve, two
    transforms each (y at f, unit weights at 2f).  Batched = one persistent
    fused launch replayed as a graph; plain = one unfused bootstrap call per
    curve (spread + batched FFT kernels), all 1024 calls replayed as one graph."""
    from tests._data import light_curves, rel_l2
    import oracle

    ncurves, M, nbs, seed, df = args.curves, 1 << 14, 100, 2024, 1.0 / 3650.0
    t, y, offs = light_curves(ncurves, 3650)
    td = torch.from_numpy(t).cuda()
    yd = torch.from_numpy(y).cuda()
    stream = torch.cuda.Stream()
    ev = HipEvents()
    sp = stream.cuda_stream
    out_y = torch.empty((ncurves * nbs, M), dtype=torch.complex64, device="cuda")
    out_w = torch.empty_like(out_y)

    def timed(graph, reps):
        e = ev.create(2)
        graph.launch(stream)  # warm
        stream.synchronize()
        ev.record(e[0], sp)
        for _ in range(reps):
            graph.launch(stream)
        ev.record(e[1], sp)
        stream.synchronize()
        ms = ev.ms(e[0], e[1]) / reps
        ev.destroy(e)
        return ms

    plan_b = cu.NFFTPlan(t=td, df=df, offsets=offs, M=M, window2=True)
    run_b = lambda: plan_b.bootstrap(yd, nbs, seed, window=True, out_y=out_y, out_w=out_w,
                                     stream=stream)
    run_b()
    stream.synchronize()
    g_b = cu.capture(run_b, stream)
    ms_b = timed(g_b, args.boot_steps)
    # accuracy of batched replicates vs the FP64 direct NDFT of the resampled points
    xs = (df * t) % 1.0
    errs = []
    for c in (0, ncurves // 2, ncurves - 1):
        n = int(offs[c + 1] - offs[c])
        d = oracle.bootstrap_indices(seed, n, nbs) + offs[c]
        for s in (0, nbs - 1):
            ref = oracle.ndft(xs[d[s]], y.astype(np.float64)[d[s]], M)
            errs.append(rel_l2(out_y[c * nbs + s].cpu().numpy(), ref))
            refw = oracle.ndft(2 * xs[d[s]], None, M)
            errs.append(rel_l2(out_w[c * nbs + s].cpu().numpy(), refw))
    keep = out_y[:nbs].cpu().numpy()
    # Lomb-Scargle periodograms of the same replicates, fused (no spectra
    # written): full power rows + peaks, and peaks only
    pw = torch.empty((ncurves * nbs, M // 2), dtype=torch.float32, device="cuda")
    pk = torch.empty(ncurves * nbs, dtype=torch.float32, device="cuda")
    kp = torch.empty(ncurves * nbs, dtype=torch.int32, device="cuda")
    run_ls = lambda: plan_b.bootstrap_ls(yd, nbs, seed, stream=stream, out_power=pw,
                                         out_peak=pk, out_kpeak=kp)
    run_pk = lambda: plan_b.bootstrap_ls(yd, nbs, seed, power=False, stream=stream,
                                         out_peak=pk, out_kpeak=kp)
    run_ls()
    stream.synchronize()
    ls_err = 0.0
    for c in (0, ncurves // 2, ncurves - 1):
        n = int(offs[c + 1] - offs[c])
        d = oracle.bootstrap_indices(seed, n, nbs) + offs[c]
        ref = oracle.lomb_scargle(t[d[0]], y.astype(np.float64)[d[0]], df, M // 2)
        ls_err = max(ls_err, float(np.max(np.abs(pw[c * nbs].cpu().numpy() - ref))))
    g_ls = cu.capture(run_ls, stream)
    ms_ls = timed(g_ls, args.boot_steps)
    g_pk = cu.capture(run_pk, stream)
    ms_pk = timed(g_pk, args.boot_steps)
    g_ls.close()
    g_pk.close()
    del pw, pk, kp

    plans = []
    for c in range(ncurves):
        sl = slice(int(offs[c]), int(offs[c + 1]))
        plans.append(cu.NFFTPlan(t=td[sl], df=df, M=M, window2=True, unfused=True))
    ys = [yd[int(offs[c]):int(offs[c + 1])] for c in range(ncurves)]

    def run_p():
        for c in range(ncurves):
            plans[c].bootstrap(ys[c], nbs, seed, window=True,
                               out_y=out_y[c * nbs:(c + 1) * nbs],
                               out_w=out_w[c * nbs:(c + 1) * nbs], stream=stream)

    run_p()
    stream.synchronize()
    agree = rel_l2(out_y[:nbs].cpu().numpy(), keep)
    g_p = cu.capture(run_p, stream)
    ms_p = timed(g_p, max(1, args.boot_steps // 2))
    g_b.close()
    g_p.close()
    for p in plans:
        p.close()
    plan_b.close()
    del out_y, out_w
    torch.cuda.empty_cache()
    npts = int(offs[-1])
    spectra = ncurves * nbs * M * 8  # one complex64 spectrum of M modes per replicate
    # compulsory: read t (FP64) and y (FP32) once, write out_y and out_w
    boot_bytes = npts * (8 + 4) + 2 * spectra
    # Lomb-Scargle: read t and y once, write M/2 FP32 powers + peak per replicate
    ls_bytes = npts * (8 + 4) + ncurves * nbs * (M // 2 * 4 + 8)
    return {
        "workload": f"{ncurves} light curves, lengths log-uniform [100, 5e4] "
                    f"({int(offs[-1])} points), M=2^14, {nbs} replicates x 2 transforms "
                    "(y at f, unit weights at 2f), FP32 + FP64 phase reduction",
        "lc_per_s_batched": ncurves / (ms_b * 1e-3),
        "lc_per_s_plain": ncurves / (ms_p * 1e-3),
        "ratio": ms_p / ms_b,
        "ms_batched": ms_b,
        "ms_plain": ms_p,
        "max_rel_err_vs_ndft": max(errs),
        "batched_vs_plain_rel_diff": agree,
        "transforms_per_s_batched": ncurves * nbs * 2 / (ms_b * 1e-3),
        "roofline": dict(hbm_roofline(boot_bytes, ms_b, copy_bw),
                         what="batched fused launch: compulsory bytes = points (t FP64 + y "
                              "FP32) read once + out_y + out_w written"),
        "lomb_scargle": {
            "what": "fused periodograms of the same replicates: power at k = 0..M/2-1 "
                    "(+ peak), or the peak only; no spectra written",
            "lc_per_s_power": ncurves / (ms_ls * 1e-3),
            "lc_per_s_peak_only": ncurves / (ms_pk * 1e-3),
            "ratio_power_vs_plain_spectra": ms_p / ms_ls,
            "ms_power": ms_ls,
            "ms_peak_only": ms_pk,
            "max_abs_err_vs_fp64_periodogram": ls_err,
            "roofline": dict(hbm_roofline(ls_bytes, ms_ls, copy_bw),
                             what="compulsory bytes = points read once + M/2 powers and the "
                                  "peak written per replicate"),
        },
    }


def bench_precision(torch, cu):
    """Mixed precision on a long baseline: N = 2e5 times over 100 years (nightly
    cadence, +-4 h jitter, 120-day seasonal gaps), M = 2^18, df = 1/(5 baseline).
    FP64 phase reduction + FP32 gridding/FFT (libcunaf) and all-FP64 (libcunad)
    against the FP64 direct NDFT, per frequency band."""
    from tests._data import long_baseline_times, rel_l2
    import oracle

    n, M = 200_000, 1 << 18
    t, y = long_baseline_times(n, 36500)
    df = 1.0 / (5.0 * (t[-1] - t[0]))
    td = torch.from_numpy(t).cuda()
    p32 = cu.NFFTPlan(t=td, df=df, M=M, dtype=torch.float32)
    p64 = cu.NFFTPlan(t=td, df=df, M=M, dtype=torch.float64)
    F32 = p32.adjoint(torch.from_numpy(y).cuda())[0].cpu().numpy()
    F64 = p64.adjoint(torch.from_numpy(y.astype(np.float64)).cuda())[0].cpu().numpy()
    # FP32 sample positions (what an FP32 x loses at this baseline)
    x32 = (((df * t) + 0.5) % 1.0 - 0.5).astype(np.float32)
    Fn = cu.NFFTPlan(torch.from_numpy(x32).cuda(), M).adjoint(
        torch.from_numpy(y).cuda())[0].cpu().numpy()
    xs = (df * t) % 1.0
    nb = 8
    k0 = (np.linspace(-M // 2, M // 2 - 64, nb)).astype(np.int64)
    ref = oracle.ndft_bands(xs, y.astype(np.float64), k0, 64)
    bands = []
    for b_, k in enumerate(k0):
        sl = slice(int(k) + M // 2, int(k) + M // 2 + 64)
        bands.append({"k0": int(k), "mixed": rel_l2(F32[sl], ref[b_]),
                      "fp64": rel_l2(F64[sl], ref[b_]), "fp32_x": rel_l2(Fn[sl], ref[b_])})
    p32.close(); p64.close()
    return {
        "workload": "N=2e5 times over 100 y (nightly cadence, +-4 h jitter, 120-d gaps), "
                    "M=2^18, df=1/(5 baseline)",
        "max_rel_err_mixed": max(b_["mixed"] for b_ in bands),
        "max_rel_err_fp64": max(b_["fp64"] for b_ in bands),
        "max_rel_err_fp32_positions": max(b_["fp32_x"] for b_ in bands),
        "bands": bands,
    }


def bench_c2_variants(x, y, torch, cu, K=50):
    """The N=1e6 clustered, M=2^20 workload through the other entry points:
    the FP64 library (accurate, m=13) and the reference's compat formula
    (cunfft_adjoint_raw_async, ng=2^21) in both libraries.  K transforms
    replayed from one graph, timed with HIP events on the launch stream."""
    from tests._data import rel_l2
    import oracle

    N, M = x.size, 1 << 20
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    ev = HipEvents()

    def graph_ms(fn):
        fn()
        stream.synchronize()
        g = cu.capture(lambda: [fn() for _ in range(K)], stream)
        g.launch(stream)
        stream.synchronize()
        e = ev.create(2)
        ev.record(e[0], sp)
        g.launch(stream)
        ev.record(e[1], sp)
        stream.synchronize()
        ms = ev.ms(e[0], e[1]) / K
        ev.destroy(e)
        g.close()
        return ms

    rng = np.random.default_rng(1)
    k0 = np.concatenate([[-M // 2, -16], rng.integers(-M // 2, M // 2 - 32, 6)]).astype(np.int64)
    ref = oracle.ndft_bands(x.astype(np.float64), y.astype(np.float64), k0, 32)
    res = {}
    xd64 = torch.from_numpy(x.astype(np.float64)).cuda()
    yd64 = torch.from_numpy(y.astype(np.float64)).cuda()
    plan = cu.NFFTPlan(xd64, M)
    out = torch.empty((1, M), dtype=torch.complex128, device="cuda")
    ms = graph_ms(lambda: plan.adjoint(yd64, out, stream=stream))
    st = ev.create(4)
    arr = (ctypes.c_void_p * 4)(*st)
    plan.L.cuna_nfft_set_profile_events(plan.h, arr, 4)
    plan.adjoint(yd64, out, stream=stream)
    stream.synchronize()
    plan.L.cuna_nfft_set_profile_events(plan.h, None, 0)
    stages = {"spread": ev.ms(st[0], st[1]), "fft pass 1": ev.ms(st[1], st[2]),
              "fft pass 2 + deconvolution": ev.ms(st[2], st[3])}
    ev.destroy(st)
    got = out[0].cpu().numpy()
    sel = np.stack([got[k + M // 2:k + M // 2 + 32] for k in k0])
    res["fp64_accurate"] = {"ms_per_transform": ms, "adj_nfft_per_s": 1e3 / ms, "m": plan.info().m,
                            "rel_err_vs_ndft": rel_l2(sel, ref), "stages_ms_one_transform": stages}
    plan.close()
    del out
    ng = 1 << 21
    for prec, xd, yd in (("fp32", torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()),
                         ("fp64", xd64, yd64)):
        fp = cu.generate_filter_properties(xd, N, ng)
        fh = torch.zeros(ng, dtype=torch.complex64 if prec == "fp32" else torch.complex128,
                         device="cuda")
        ms = graph_ms(lambda: cu.cunfft_adjoint_raw_async(xd, yd, fh, N, ng, fp, stream=stream))
        res[prec + "_compat_raw_async"] = {"ms_per_transform": ms, "adj_nfft_per_s": 1e3 / ms}
        fp.free()
        del fh
    res["workload"] = ("N=1e6 clustered, M=2^20 (accurate, sigma=2) / ng=2^21 (compat, the "
                       "reference's formula through cunfft_adjoint_raw_async); plan reused")
    return res


def bench_compat_batch(torch, cu, nlc=256, K=5):
    """The reference formula over many curves: nlc light curves of
    log-uniform length in [100, 5e4], each with ng = the next power of two
    above its length, through one cunfft_adjoint_raw_batched_async call
    (cached multi-curve plan: one spread + one FFT launch per distinct ng)
    against nlc cunfft_adjoint_raw_async calls; both replayed from graphs."""
    from tests._data import rel_l2
    import oracle

    rng = np.random.default_rng(77)
    n = np.exp(rng.uniform(np.log(100), np.log(5e4), nlc)).astype(np.int64)
    ng = [1 << int(np.ceil(np.log2(v + 1))) for v in n]
    n = [int(v) for v in n]
    off = np.concatenate([[0], np.cumsum(n)])
    x = rng.uniform(-0.5, 0.5, off[-1]).astype(np.float32)
    y = rng.standard_normal(off[-1]).astype(np.float32)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    idist = max(ng)
    fps = [cu.generate_filter_properties(xd[off[i]:off[i + 1]], n[i], ng[i]) for i in range(nlc)]
    fh = torch.zeros(nlc * idist, dtype=torch.complex64, device="cuda")
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    ev = HipEvents()

    def graph_ms(fn):
        fn()
        stream.synchronize()
        g = cu.capture(fn, stream)
        g.launch(stream)
        stream.synchronize()
        e = ev.create(2)
        ev.record(e[0], sp)
        for _ in range(K):
            g.launch(stream)
        ev.record(e[1], sp)
        stream.synchronize()
        ms = ev.ms(e[0], e[1]) / K
        ev.destroy(e)
        g.close()
        return ms

    batched = lambda: cu.cunfft_adjoint_raw_batched_async(xd, yd, fh, n, ng, fps, stream=stream)
    ms_b = graph_ms(batched)
    got = fh.cpu().numpy().reshape(nlc, idist)
    err = 0.0
    for i in (0, nlc // 2, nlc - 1):
        sl = slice(off[i], off[i + 1])
        err = max(err, rel_l2(got[i, :ng[i]], oracle.compat_adjoint(x[sl], y[sl], ng[i], False)))
    views = [(xd[off[i]:off[i + 1]], yd[off[i]:off[i + 1]], fh[i * idist:i * idist + ng[i]])
             for i in range(nlc)]

    def loop():
        for i, (xv, yv, fv) in enumerate(views):
            cu.cunfft_adjoint_raw_async(xv, yv, fv, n[i], ng[i], fps[i], stream=stream)

    ms_l = graph_ms(loop)
    for f in fps:
        f.free()
    cu.batch_cache_clear("f")
    return {
        "workload": f"{nlc} curves, n log-uniform in [100, 5e4] ({off[-1]} points), ng = next "
                    f"power of two above n ({len(set(ng))} distinct), FP32 compat formula",
        "batched_ms": ms_b, "per_curve_calls_ms": ms_l, "speedup": ms_l / ms_b,
        "curves_per_s_batched": nlc / (ms_b * 1e-3), "max_rel_err_vs_oracle": err,
    }


def bench_differentiation(torch, cu, steps=20):
    """64 curves x N = 2e4, M = 2^16, L = sum_k w_k |F_k|^2: forward (adjoint)
    and backward (d/dy and d/dx) times of the batch, each replayed as a graph."""
    from tests._data import rel_l2
    import oracle

    B, n, M = 64, 20_000, 1 << 16
    rng = np.random.default_rng(5)
    x = rng.uniform(-0.5, 0.5, (B, n)).astype(np.float32)
    y = rng.standard_normal((B, n)).astype(np.float32)
    w = rng.uniform(0, 1, M).astype(np.float32)
    xd = torch.from_numpy(x.reshape(-1)).cuda()
    yd = torch.from_numpy(y.reshape(-1)).cuda()
    wd = torch.from_numpy(w).cuda()
    offs = np.arange(B + 1, dtype=np.int32) * n
    plan = cu.NFFTPlan(xd, M, offsets=offs)
    stream = torch.cuda.Stream()
    out = torch.empty((B, M), dtype=torch.complex64, device="cuda")
    g = torch.empty_like(out)
    dy = torch.empty(B * n, dtype=torch.float32, device="cuda")
    dx = torch.empty_like(dy)
    L = plan.L

    def fwd():
        plan.adjoint(yd, out, stream=stream)

    def bwd():
        st = L.cuna_nfft_backward(plan.h, yd.data_ptr(), g.data_ptr(), dy.data_ptr(),
                                  dx.data_ptr(), stream.cuda_stream)
        assert st == 0

    fwd()
    stream.synchronize()
    g.copy_(2 * wd * out)
    bwd()
    stream.synchronize()
    ev = HipEvents()
    res = {}
    for name, fn in (("forward", fwd), ("backward", bwd)):
        gr = cu.capture(lambda: [fn() for _ in range(steps)], stream)
        gr.launch(stream)
        stream.synchronize()
        e = ev.create(2)
        ev.record(e[0], stream.cuda_stream)
        gr.launch(stream)
        ev.record(e[1], stream.cuda_stream)
        stream.synchronize()
        res[name] = ev.ms(e[0], e[1]) / steps
        ev.destroy(e)
        gr.close()
    # gradient of curve 0 vs the analytic FP64 NDFT gradient
    F0 = oracle.ndft(x[0].astype(np.float64), y[0].astype(np.float64), M)
    ady, adx = oracle.ndft_grad(x[0].astype(np.float64), y[0].astype(np.float64), M,
                                2 * w.astype(np.float64) * F0)
    gy = dy[:n].cpu().numpy()
    gx = dx[:n].cpu().numpy()
    plan.close()
    return {
        "workload": "64 curves x N=2e4, x~U, y~N(0,1), M=2^16, L=sum w_k|F_k|^2, FP32",
        "forward_ms": res["forward"],
        "backward_ms": res["backward"],
        "backward_over_forward": res["backward"] / res["forward"],
        "grad_y_rel_err_vs_ndft": rel_l2(gy, ady),
        "grad_x_rel_err_vs_ndft": rel_l2(gx, adx),
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-bootstrap", action="store_true")
    ap.add_argument("--curves", type=int, default=1024)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the extra workloads (FP64 and compat variants of the N=1e6 "
                         "transform, batched compat curves, precision, differentiation)")
    ap.add_argument("--extras", default="c2_variants,compat_batch,precision,differentiation",
                    help="comma-separated subset of the extra workloads to run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one device copy of exactly 256 MiB (a known-size kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--boot-steps", type=int, default=4)
    ap.add_argument("--no-multi-rhs", action="store_true",
                    help="skip the several-value-vectors measurement (keeps a profile of the "
                         "headline kernels to one-vector calls)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: it measures the single-GPU kernels. Over several "
              "GPUs a transform would shard by points (one reduce-scatter of the spectra) and "
              "batches by curve (no exchange); that path is not built",
              file=sys.stderr)
        return 2

    import torch
    import cunfft_adjoint_amd as cu

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    cu.load("f")
    if args.calibrate_copy:
        a = torch.empty(1 << 26, dtype=torch.float32, device="cuda")  # 256 MiB
        c = torch.empty_like(a)
        c.copy_(a)
        torch.cuda.synchronize()
        del a, c
    r = bench_c2(args, torch, cu)
    bw = copy_bandwidth(torch)
    res = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "adj-NFFT/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain adjoint NFFT, FP32 library, accurate mode: N=1e6 clustered "
                        "points (50 clumps of width 0.005 + 20% uniform), M=2^20, sigma=2 "
                        "(G=2^21), m=8; plan reused",
            "N": 1_000_000, "M": r["M"], "G": r["info"]["G"], "m": r["info"]["m"],
            "tile_log2": r["info"]["tile_log2"],
        },
        "rel_err_vs_ndft": r["rel_err_vs_ndft"],
        "plan_ms": r["plan_ms"],
        "plan_warm": r["plan_warm"],
        "eager_ms_per_step": r["eager_ms_per_step"],
        "graph_ms_per_step": r["graph_ms_per_step"],
        "timing_host_bound": r["timing_host_bound"],
        "stages_ms": r["stages_ms"],
        "roofline": r["roofline"],
        "pipeline": r["pipeline"],
        "multi_rhs": r["multi_rhs"],
        "copy_bw_GBps": bw / 1e9,
        "cpu_baseline": None,
    }
    res["roofline"]["frac_of_copy_bw"] = res["roofline"]["achieved"] * 1e9 / bw
    res["pipeline"]["frac_of_copy_bw"] = res["pipeline"]["achieved_GBps"] * 1e9 / bw
    if not args.no_bootstrap:
        res["bootstrap"] = bench_bootstrap(args, torch, cu, bw)
    if not args.no_extras:
        cu.load("d")
        extras = {
            "c2_variants": lambda: bench_c2_variants(r["x"], r["y"], torch, cu),
            "compat_batch": lambda: bench_compat_batch(torch, cu),
            "precision": lambda: bench_precision(torch, cu),
            "differentiation": lambda: bench_differentiation(torch, cu),
        }
        for name in args.extras.split(","):
            if name:
                if name not in extras:
                    print(f"unknown extra workload {name!r}", file=sys.stderr)
                    return 2
                res[name] = extras[name]()
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(r["x"], r["y"], r["M"], args.cpu_seconds)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
