#!/usr/bin/env python
"""Benchmark: GLS periodogram evaluations/s on one MI355X.

Workload (one "step"): the survey batch -- 5,000 light curves with N
log-uniform in [50, 5000], baselines U[30, 3650] d with seasonal gaps and
nightly jitter, linear grids df = 1/(5 T) up to 24 /d (1.07e9 curve x
frequency evaluations) -- run through the batched GPU path with inputs already
resident in HBM.  Also measured: the plain one-curve-per-call path of the
drop-in API on a sample, per-stage kernel times (HIP events on the launch
stream), the FP64 CPU port on the host (cpu_baseline), and the parity of the
GPU periodograms against that port's own outputs on its first sampled curves.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
TOL = 1e-4


def _cpu_worker(args):
    # oracle FP64 NFFT-GLS of one curve (the reference-style CPU path); the
    # periodogram is handed back for the curves kept as the parity sample
    from oracle import nfft as onfft
    t, y, dy, freqs, keep = args
    P = onfft.gls_nfft(t, y, dy, freqs, sigma=2.0, m=8)
    return len(freqs), (P if keep else None)


def long_curve_oracle():
    """FP64 oracle periodogram of the long curve (200,000 points, ~1.8M
    frequencies; Gaussian sigma = 2, m = 10 NFFT, |dP| ~1e-9): the reference
    of the precision study, computed by the CPU-baseline leg."""
    from cuvarbase_amd import synthetic
    from oracle import nfft as onfft

    data, freqs = synthetic.long()
    t, y, dy = data[0]
    return onfft.gls_nfft(t, y, dy, freqs[0], sigma=2.0, m=10)


def cpu_baseline(data, freqs, target_s=12.0, max_curves=400, seed=7, keep=8, single_s=4.0):
    """Time the FP64 oracle port (numpy NFFT + GLS) on a bounded sample.

    Returns the baseline record and {curve index: oracle periodogram} for the
    first ``keep`` curves of the sample, against which the GPU results are
    compared later (the oracle runs only here)."""
    import multiprocessing as mp

    cores = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(data))
    # calibrate on one curve, then size the sample to ~target_s of CPU work
    i0 = int(order[0])
    t0 = time.perf_counter()
    _cpu_worker((*data[i0], freqs[i0], False))
    per_eval = (time.perf_counter() - t0) / len(freqs[i0])
    budget_evals = target_s / max(per_eval, 1e-12)
    sample, ev = [], 0
    for i in order[1:]:
        if ev >= budget_evals or len(sample) >= max_curves:
            break
        sample.append(int(i))
        ev += len(freqs[i])
    ctx = mp.get_context("spawn")
    with ctx.Pool(cores) as pool:
        t0 = time.perf_counter()
        res = pool.map(_cpu_worker, [(*data[i], freqs[i], j < keep)
                                     for j, i in enumerate(sample)], chunksize=1)
        wall = time.perf_counter() - t0
    done = sum(r[0] for r in res)
    kept = {i: r[1] for i, r in zip(sample, res) if r[1] is not None}
    # one core: the same port in this process over the sample's first curves,
    # about single_s seconds of work
    t0 = time.perf_counter()
    one, ncur = 0, 0
    for i in sample:
        if time.perf_counter() - t0 >= single_s:
            break
        one += _cpu_worker((*data[i], freqs[i], False))[0]
        ncur += 1
    one_wall = time.perf_counter() - t0
    return {
        "value": done / wall,
        "unit": "evals/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{len(sample)} randomly chosen survey curves ({done:.3g} evals), FP64 "
                  f"NumPy NFFT (sigma=2, m=8) + GLS, one curve per task on a "
                  f"{cores}-process pool",
        "wall_s": wall,
        "single_core": {"value": one / one_wall, "unit": "evals/s", "cores": 1,
                        "sample": f"the first {ncur} curves of that sample ({one:.3g} evals) "
                                  "in one process"},
    }, kept


def profiled_traffic(kernel):
    """HBM bytes per launch of ``kernel`` (read + write, scaled by the copy
    kernel) from the committed rocprofv3 PMC summary of this workload, or
    None if the summary does not list it."""
    import re

    path = os.path.join(HERE, "profiles", "survey5000_passes_kernel_stats.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"traffic_bytes_per_launch: (.+?) = (\d+)\s*$", line)
        if m and m.group(1) == kernel:
            return float(m.group(2))
    return None


def measured_copy_gbs(stream, mib=2048, reps=10):
    """Device-to-device copy rate (read + write bytes per second) of the HIP
    runtime's copy kernel on ``stream``: a measured HBM ceiling next to the
    8 TB/s published peak, and the known-size kernel that scales the HBM
    byte counters of a profiler run."""
    from cuvarbase_amd import _device, _lib
    import ctypes as C

    nbytes = mib << 20
    a = _device.DeviceArray(nbytes, np.uint8)
    b = _device.DeviceArray(nbytes, np.uint8)
    h = _device._handle(stream)
    _lib.call("gls_memcpy_d2d_async", b.ptr, a.ptr, C.c_size_t(nbytes), h)
    e0, e1 = _device.Event(), _device.Event()
    e0.record(stream)
    for _ in range(reps):
        _lib.call("gls_memcpy_d2d_async", b.ptr, a.ptr, C.c_size_t(nbytes), h)
    e1.record(stream)
    e1.synchronize()
    ms = e1.time_since(e0) / reps
    a.free()
    b.free()
    return 2.0 * nbytes / (ms / 1e3) / 1e9


def _parity(gpu, oracle):
    """max |dP| and best-index agreement of {curve: powers} against the
    oracle's periodograms of the same curves."""
    if not oracle:
        return None
    worst, same = 0.0, 0
    for i, Po in oracle.items():
        Pg = np.asarray(gpu[i], dtype=np.float64)
        worst = max(worst, float(np.max(np.abs(Pg - Po))))
        same += int(np.argmax(Pg) == np.argmax(Po))
    return {"max_abs_dP": worst, "best_period_match": f"{same}/{len(oracle)}",
            "tolerance": TOL,
            "oracle": "FP64 NumPy NFFT sigma=2 m=8 + GLS (the CPU-baseline leg's own outputs "
                      "on its first sampled curves)"}


def precision_study(data, freqs, batch, oracle_P, sigma, m, win, stream):
    """max |dP| and best period against the FP64 oracle of the long curve:
    FP32 with FP64 time/cell/phase reduction (``batch``, the default), and
    FP32 everywhere (the reference's FP32 time arithmetic, same kernels)."""
    from cuvarbase_amd import _lib, engine

    f = freqs[0]
    kbest = int(np.argmax(oracle_P))
    out = {"oracle": "FP64 NumPy NFFT (Gaussian sigma=2, m=10) + FP64 GLS, computed in the "
                     "CPU-baseline leg", "oracle_best_period_d": float(1 / f[kbest])}
    fp32 = engine.LombBatch(data, freqs, sigma=sigma, m=m, precision=_lib.PRECISION_FP32, **win)
    for name, b in (("fp32_with_fp64_phase", batch), ("fp32_everywhere", fp32)):
        b.upload(stream=stream)
        b.execute(stream=stream)
        P = b.split(b.download_async(stream=stream))[0]
        stream.synchronize()
        P = np.asarray(P, dtype=np.float64)
        k = int(np.argmax(P))
        out[name] = {"max_abs_dP": float(np.max(np.abs(P - oracle_P))),
                     "best_period_d": float(1 / f[k]), "same_best_index": k == kbest}
    return out


def end_to_end(batch, stream, reps=3, k=5):
    """Batch throughput with the inputs starting and the results ending in
    page-locked host memory: H2D of t, y, dy, the execute, and D2H of every
    power (or of the k highest peaks per curve).  HIP events on ``stream``."""
    import ctypes as C

    from cuvarbase_amd import _device, _lib

    h = _device._handle(stream)
    t, a, b = batch._host
    hin = [_device.pinned_empty(len(x), x.dtype) for x in (t, a, b)]
    for dst, src in zip(hin, (t, a, b)):
        dst[:] = src
    hout = _device.pinned_empty(max(batch.total_out, 1), np.float32)
    ncur = batch.ncurves
    tp = _device.DeviceArray(ncur * k, np.float32)
    ti = _device.DeviceArray(ncur * k, np.int32)
    htop = _device.pinned_empty(2 * ncur * k, np.float32)
    e0, e1 = _device.Event(), _device.Event()

    def upload():
        for dev, src in zip((batch.d_t, batch.d_a, batch.d_b), hin):
            _lib.call("gls_memcpy_h2d_async", dev.ptr, src.ctypes.data, C.c_size_t(src.nbytes), h)

    def full():
        upload()
        batch.execute(stream=stream)
        _lib.call("gls_memcpy_d2h_async", hout.ctypes.data, batch.d_power.ptr,
                  C.c_size_t(4 * batch.total_out), h)

    def peaks():
        upload()
        batch.execute(stream=stream)
        _lib.call("gls_topk", batch.plan.handle, batch.d_ws.ptr, batch.d_power.ptr, None,
                  C.c_int32(k), tp.ptr, ti.ptr, h)
        _lib.call("gls_memcpy_d2h_async", htop.ctypes.data, tp.ptr, C.c_size_t(4 * ncur * k), h)
        _lib.call("gls_memcpy_d2h_async", htop.ctypes.data + 4 * ncur * k, ti.ptr,
                  C.c_size_t(4 * ncur * k), h)

    out = {}
    for name, fn in (("all_powers", full), ("top5_peaks", peaks)):
        fn()
        stream.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        ms = e1.time_since(e0) / reps
        out[name] = {"ms": ms, "evals_per_s": batch.total_out / (ms / 1e3),
                     "d2h_bytes": 4 * batch.total_out if name == "all_powers" else 8 * ncur * k}
    out["h2d_bytes"] = int(sum(x.nbytes for x in hin))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ncurves", type=int, default=5000)
    ap.add_argument("--plain-sample", type=int, default=300)
    ap.add_argument("--parity-sample", type=int, default=8,
                    help="curves of the CPU-baseline sample whose oracle periodograms "
                         "are kept for the parity check")
    ap.add_argument("--shapes", type=int, default=1,
                    help="time the single-curve and long-curve shapes (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sigma", type=float, default=None,
                    help="NFFT oversampling (default: tuned for |dP| <= 1e-4)")
    ap.add_argument("--m", type=int, default=None)
    ap.add_argument("--window", type=int, default=None,
                    help="NFFT window: 0 Gaussian, 1 exponential of semicircle (default: tuned)")
    ap.add_argument("--stream-requests", type=int, default=5000,
                    help="requests per arrival rate in the serving leg, at most 10 s of "
                         "arrivals per rate (0: skip)")
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0)
    ap.add_argument("--short-curves", type=int, default=100_000,
                    help="curves in the short-curve top-5 leg (0: skip)")
    ap.add_argument("--fork-steps", type=int, default=5,
                    help="extra steps timed with concurrent size buckets (0: skip)")
    ap.add_argument("--e2e", type=int, default=1,
                    help="time the end-to-end (host memory to host memory) batch (0: skip)")
    ap.add_argument("--spread", type=int, default=2, choices=(0, 1, 2),
                    help="batched path's pass-1 accumulation: 0 FP32 atomics, 1 fixed point "
                         "(bitwise reproducible), 2 per curve by point density (the "
                         "default); the server always runs fixed point")
    ap.add_argument("--grid-chunk-mib", type=int, default=0,
                    help="NFFT grid chunk size in MiB (0: no chunking)")
    ap.add_argument("--calib-copy-mib", type=int, default=2048,
                    help="size of the device-to-device copies timed before the steps "
                         "(measured copy rate; scales HBM counters of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the GLS path "
              "shards curves across GPUs with no exchange, but this build measures one "
              "GPU's share.", file=sys.stderr)
        sys.exit(2)

    from cuvarbase_amd import synthetic, tuning

    cfg = tuning.choose_nfft_config(TOL)
    if args.sigma is not None or args.m is not None or args.window is not None:
        cfg = dict(sigma=args.sigma if args.sigma is not None else cfg["sigma"],
                   m=args.m if args.m is not None else cfg["m"],
                   window=args.window if args.window is not None else cfg["window"],
                   window_beta=0.0)
    sigma, m = cfg["sigma"], cfg["m"]
    win = dict(window=cfg["window"], window_beta=cfg["window_beta"])
    data, freqs = synthetic.survey(args.ncurves)
    total_evals = int(sum(len(f) for f in freqs))

    cpu, oracle_P = (None, {}) if args.no_cpu else cpu_baseline(data, freqs,
                                                                 keep=args.parity_sample)
    long_P = None if (args.no_cpu or not args.shapes) else long_curve_oracle()

    from cuvarbase_amd import engine, _device, _lib
    from cuvarbase_amd.lombscargle import LombScargleAsyncProcess

    chunk = args.grid_chunk_mib << 20 if args.grid_chunk_mib > 0 else 0
    batch = engine.LombBatch(data, freqs, sigma=sigma, m=m, grid_chunk_bytes=chunk,
                             spread=args.spread, **win)
    s = _device.Stream()
    batch.upload(stream=s)
    s.synchronize()
    for _ in range(args.warmup):
        batch.execute(stream=s)
    s.synchronize()
    copy_gbs = measured_copy_gbs(s, mib=max(args.calib_copy_mib, 256))

    # timed steps (default launch order: size buckets in series on one stream)
    e0, e1 = _device.Event(), _device.Event()
    step_ms = []
    _device.synchronize()
    t_wall = time.perf_counter()
    for _ in range(args.steps):
        e0.record(s)
        batch.execute(stream=s)
        e1.record(s)
        e1.synchronize()
        step_ms.append(e1.time_since(e0))
    _device.synchronize()
    wall = time.perf_counter() - t_wall
    ms_step = float(np.mean(step_ms))
    value = total_evals / (ms_step / 1e3)

    # instrumented steps (after the timed ones): HIP events between launches
    # on one stream give each stage's and kernel's duration, launches in
    # series so that each is measured alone
    batch.plan.enable_timing(True)
    stage_acc = {}
    n_instr = max(1, min(args.steps, 3))
    for _ in range(n_instr):
        batch.execute(stream=s)
        s.synchronize()
        for i, (name, p, ms, by) in enumerate(batch.plan.stage_times()):
            a = stage_acc.setdefault((i, name), [0, 0.0, 0.0])
            a[0] += 1
            a[1] += ms
            a[2] += by
    batch.plan.enable_timing(False)
    serial_ms = sum(v[1] for v in stage_acc.values()) / n_instr

    # the optional concurrent launch order (gls_plan_set_fork: size buckets
    # spread over the library's streams, so one bucket's VALU-bound pass 1
    # can overlap another's HBM-bound pass 2), timed like the steps above;
    # reported beside the headline, which keeps the serial default
    fork_ms = None
    if args.fork_steps > 0:
        batch.plan.set_fork(True)
        batch.execute(stream=s)
        s.synchronize()
        e0.record(s)
        for _ in range(args.fork_steps):
            batch.execute(stream=s)
        e1.record(s)
        e1.synchronize()
        fork_ms = e1.time_since(e0) / args.fork_steps
        batch.plan.set_fork(False)

    ncurves = batch.ncurves
    ws_bytes = batch.plan.workspace_bytes
    geo = [batch.plan.geometry(i) for i in range(ncurves)]
    n_fixed = sum(batch.plan.spread(i) == _lib.SPREAD_FIXED for i in range(ncurves))

    # dominant kernel: the kernel template (pass-1 column length n1, tile c1
    # and accumulation, or pass-2 row length n2 and tile cr) with the most
    # time over the instrumented steps
    by_kernel, by_kind = {}, {}
    for (i, name), (c, ms, by) in stage_acc.items():
        n1, p2, c1, cr, sp, nts = batch.plan.stage_bucket(i)
        if name == "pass1":
            fx = "true" if sp == _lib.SPREAD_FIXED else "false"
            key = f"pass1_kernel<{n1}, {c1}, {fx}, {'true' if nts else 'false'}>"
        elif name == "pass2":
            key = f"pass2_kernel<{p2}, {cr}, 0>"
        else:
            key = name
        for d, k in ((by_kernel, key), (by_kind, name)):
            a = d.setdefault(k, [0, 0.0, 0.0])
            a[0] += c
            a[1] += ms
            a[2] += by
    dom = max(by_kernel, key=lambda k: by_kernel[k][1])
    c, ms, by = by_kernel[dom]
    achieved = (by / c) / (ms / c / 1e3) / 1e9
    traffic = profiled_traffic(dom)

    # parity on the curves whose FP64 oracle periodograms the CPU-baseline leg
    # kept (randomly chosen survey curves)
    flat = batch.download_async(stream=s)
    s.synchronize()
    P = batch.split(flat)
    parity = _parity({i: P[i] for i in oracle_P}, oracle_P)

    # end to end: the same batch from page-locked host memory to page-locked
    # host memory (H2D of t, y, dy + execute + D2H of every power, or of the
    # top-5 peaks per curve), on the execute stream
    e2e = end_to_end(batch, s, reps=max(1, min(args.steps, 3))) if args.e2e else None

    # plain path: the drop-in API, one curve per run() + finish(), at the same
    # NFFT configuration as the batched path
    plain_kw = dict(sigma=sigma, m=m, nfft_window="es" if cfg["window"] == 1 else "gaussian",
                    window_beta=cfg["window_beta"])
    proc = LombScargleAsyncProcess(**plain_kw)
    nplain = min(args.plain_sample, len(data))
    proc.run([data[0]], freqs=[freqs[0]])
    proc.finish()
    t0 = time.perf_counter()
    ev = 0
    for i in range(nplain):
        proc.run([data[i]], freqs=[freqs[i]])
        proc.finish()
        ev += len(freqs[i])
    plain = ev / (time.perf_counter() - t0) if nplain else None
    plain_P = {}
    for i in oracle_P:
        (_, p), = proc.run([data[i]], freqs=[freqs[i]])
        proc.finish()
        plain_P[i] = np.array(p)
    plain_parity = _parity(plain_P, oracle_P)

    # served path: continuous batching of host-resident requests arriving as a
    # Poisson stream (PCIe transfers included), and the plain path on the same
    # requests, one per run() + finish()
    serving = None
    if args.stream_requests > 0:
        del batch, flat, P
        _device.synchronize()
        from cuvarbase_amd.serve import GLSServer

        arr, sdata, sfreqs = synthetic.stream_requests(1.0, nreq=args.stream_requests)
        srv = GLSServer(hbm_budget=args.hbm_budget_gb * 1e9, slots=3, sigma=sigma, m=m,
                        zero_copy=True, max_batch_points=8_000_000, **win)
        serving = {"requests_per_rate": args.stream_requests,
                   "rule": "per rate: the requested count or 10 s of arrivals, whichever "
                           "is fewer", "hbm_budget_gb": args.hbm_budget_gb, "rates": {}}
        for rate in (100.0, 1000.0, 10000.0, float("inf")):
            nreq = args.stream_requests if math.isinf(rate) else \
                min(args.stream_requests, int(rate * 10))
            a = np.zeros(nreq) if math.isinf(rate) else arr[:nreq] / rate
            st = srv.serve_stream(a, sdata[:nreq], sfreqs[:nreq], release=True)
            srv.completed.clear()
            serving["rates"]["burst" if math.isinf(rate) else f"{rate:g}/s"] = {
                k: st[k] for k in ("requests", "evals", "wall_s", "evals_per_s", "p50_ms",
                                   "p99_ms", "batches", "launch_host_ms", "slow_launches")}
        t0 = time.perf_counter()
        ev = 0
        nps = min(100, len(sdata))
        for i in range(nps):
            proc.run([sdata[i]], freqs=[sfreqs[i]])
            proc.finish()
            ev += len(sfreqs[i])
        serving["plain_evals_per_s"] = ev / (time.perf_counter() - t0)
        serving["plain_config"] = dict(plain_kw, path="LombScargleAsyncProcess.run, one request "
                                       "per call + finish(), PCIe included")
        best = max(v["evals_per_s"] for v in serving["rates"].values())
        serving["served_vs_plain"] = best / serving["plain_evals_per_s"]
        for rate in serving["rates"].values():
            rate["vs_plain"] = rate["evals_per_s"] / serving["plain_evals_per_s"]
        del srv
        _device.synchronize()

    # short-curve leg: 100,000 curves of 20-300 points on one shared grid of
    # 20,000 frequencies, top-5 peaks per curve on the device (inputs resident)
    short_leg = None
    if args.short_curves > 0:
        import ctypes as C
        from cuvarbase_amd import _lib

        if "batch" in locals():
            del batch
        _device.synchronize()
        shd, shf = synthetic.short(args.short_curves)
        sb = engine.LombBatch(shd, shf, sigma=sigma, m=m, **win)
        sb.upload(stream=s)
        k = 5
        tp = _device.DeviceArray(sb.ncurves * k, np.float32)
        ti = _device.DeviceArray(sb.ncurves * k, np.int32)

        def short_step():
            sb.execute(stream=s)
            _lib.call("gls_topk", sb.plan.handle, sb.d_ws.ptr, sb.d_power.ptr, None,
                      C.c_int32(k), tp.ptr, ti.ptr, _device._handle(s))

        short_step()
        s.synchronize()
        e0.record(s)
        for _ in range(3):
            short_step()
        e1.record(s)
        e1.synchronize()
        sms = e1.time_since(e0) / 3
        short_leg = {"curves": sb.ncurves, "evals": int(sb.total_out), "ms": sms,
                     "evals_per_s": sb.total_out / (sms / 1e3), "top_k": k,
                     "path": "engine, inputs resident, one execute + top-5 on the device",
                     "workload": "N~U{20..300}, t~U[0,90] d, 20,000 freqs in [0.01, 40] /d"}
        del sb, tp, ti
        _device.synchronize()
        # the same 100,000 requests through the server as one burst: batches
        # on the shared grid replay captured executes, top-5 per request
        from cuvarbase_amd.serve import GLSServer

        srv5 = GLSServer(hbm_budget=min(args.hbm_budget_gb, 100.0) * 1e9, slots=3,
                         sigma=sigma, m=m,
                         zero_copy=True, max_batch_points=8_000_000, **win)
        st5 = srv5.serve_stream(np.zeros(len(shd)), shd, shf, release=True, k=k)
        short_leg["served"] = {key: st5[key] for key in (
            "requests", "evals", "wall_s", "evals_per_s", "p50_ms", "p99_ms", "batches",
            "graph_captures", "graph_replays")}
        short_leg["served"]["path"] = ("GLSServer burst, PCIe included, k=5 peaks per "
                                       "request, captured executes replayed per shape")
        del srv5, st5
        _device.synchronize()

    # one-curve shapes through the same engine, inputs resident: the single
    # curve (1,000 points x 10,000 frequencies) and the long curve (200,000
    # points over 3650 d, ~1.8e6 frequencies to 100 /d); eager executes and
    # replays of the captured execute
    shapes = None
    if args.shapes:
        shapes = {}
        for name, (cd, cf) in (("single_1000pts_10000freqs", synthetic.single()),
                               ("long_200000pts_1.8Mfreqs", synthetic.long())):
            cb = engine.LombBatch(cd, cf, sigma=sigma, m=m, **win)
            cb.upload(stream=s)
            cb.execute(stream=s)
            s.synchronize()
            e0.record(s)
            for _ in range(10):
                cb.execute(stream=s)
            e1.record(s)
            e1.synchronize()
            cms = e1.time_since(e0) / 10
            g = cb.capture(s)
            g.launch(s)
            s.synchronize()
            e0.record(s)
            for _ in range(10):
                g.launch(s)
            e1.record(s)
            e1.synchronize()
            gms = e1.time_since(e0) / 10
            g.destroy()
            shapes[name] = {"points": len(cd[0][0]), "evals": int(cb.total_out),
                            "ms": cms, "evals_per_s": cb.total_out / (cms / 1e3),
                            "graph_ms": gms}
            if name.startswith("long") and long_P is not None:
                shapes[name]["precision_study"] = precision_study(cd, cf, cb, long_P, sigma,
                                                                  m, win, s)
            del cb
        _device.synchronize()

    # pipeline-level traffic at the executed grid sizes: the fused pipeline's
    # minimum (inputs, one write and one read of the three pass-1 grids per
    # curve, powers) over the step time
    nobs = np.array([len(d[0]) for d in data], dtype=np.float64)
    nfa = np.array([len(f) for f in freqs], dtype=np.float64)
    cells = np.array([3.0 * g[0] for g in geo])
    pipe_bytes = float(np.sum(24 * nobs + 16 * cells + 4 * nfa))

    latency = None
    if serving is not None:
        r = serving["rates"]["10000/s"]
        latency = {"p50": r["p50_ms"], "p99": r["p99_ms"],
                   "source": "served requests at 10,000 arrivals/s (serving.rates)"}

    out = {
        "metric": "LS evals/s (curve x freq), batched GPU path, survey batch",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"survey-{ncurves}: N log-uniform [50,5000], baselines "
                        "U[30,3650] d, seasonal gaps, df=1/(5T), fmax=24/d",
            "curves": ncurves,
            "evals_per_step": total_evals,
            "sigma": sigma,
            "m": m,
            "window": "exponential of semicircle" if cfg["window"] == 1 else "gaussian",
            "window_beta": cfg["window_beta"],
            "grid_cells_per_step": int(3 * sum(g[0] for g in geo)),
            "grid_chunk_bytes": int(chunk),
            "spread": {0: "fp32 atomics", 1: "fixed point",
                       2: f"per curve: fixed point for {n_fixed} curves of at least one "
                          "point per 80 cells on columns of 64+ cells, fp32 atomics for the "
                          "others"}[args.spread],
            "workspace_bytes": int(ws_bytes),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "launches_per_step": c // n_instr,
            "avg_launch_ms": ms / c,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "profiles/survey5000_passes_kernel_stats.md (rocprofv3 --pmc "
                              "FETCH_SIZE, WRITE_SIZE; scaled by the copy kernel)"
                              if traffic is not None else None,
            "algorithmic_bytes_per_launch": by / c,
            "measured_copy_GBs": copy_gbs,
            "frac_of_measured_copy": achieved / copy_gbs,
        },
        "pipeline_traffic": {
            "definition": "24 B per observation (t, y, dy) + 16 B per cell of the three "
                          "grids at the executed sizes (pass-1 write, pass-2 read) + 4 B "
                          "per power, over the step time",
            "bytes_per_step": pipe_bytes,
            "achieved_GBs": pipe_bytes / (ms_step / 1e3) / 1e9,
            "frac": pipe_bytes / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS,
        },
        "stages_ms_serial": {k: v[1] / n_instr for k, v in by_kind.items()},
        "kernels_ms_serial": {k: v[1] / n_instr for k, v in
                                sorted(by_kernel.items(), key=lambda kv: -kv[1][1])[:8]},
        "plain_evals_per_s": plain,
        "plain_config": plain_kw,
        "plain_parity": plain_parity,
        "device_resident_batched_vs_plain": (value / plain) if plain else None,
        "end_to_end": e2e,
        "latency_ms": latency,
        "serial_launch_order_ms": serial_ms,
        "concurrent_buckets": None if fork_ms is None else {
            "ms_per_step": fork_ms, "evals_per_s": total_evals / (fork_ms / 1e3),
            "note": "gls_plan_set_fork: size buckets on concurrent streams (opt-in); "
                    "the headline value keeps the serial default"},
        "parity": parity,
        "serving": serving,
        "short_top5": short_leg,
        "one_curve_shapes": shapes,
        "cpu_baseline": cpu,
        "wall_s_timed": wall,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
