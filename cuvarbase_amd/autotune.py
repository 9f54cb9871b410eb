"""Measured tiles and costs of the NFFT passes per size bucket.

    python -m cuvarbase_amd.build --sweep      # once, on the build machine
    python -m cuvarbase_amd.autotune --out cuvarbase_amd/tiles_gfx950.json

A size bucket is a four-step split n = n1 * 2^p2 of the grid.  Its tiles are
c1, the grid columns one pass-1 workgroup spreads and transforms, and cr, the
grid rows one pass-2 workgroup transforms (csrc/tiles.hpp).  This script runs
on the device the table is for, with the tile-sweep library (a kernel for every
candidate tile).  For every bucket of the catalogue (1024 <= n <= 2^24) it
plans a synthetic batch of curves with nf = n / sigma frequencies at the tuned
NFFT parameters (sigma = 1.35, m = 5, ES window), forced onto that bucket, of
about CELLS grid cells per grid, whose point counts follow the survey
workload's (log-uniform over POINTS_MIN..POINTS_MAX), spread as the batched
path's default does (fixed point for the curves of at least one point per 80
cells on columns of at least 64 cells, FP32 atomics for the others).  It times pass 1 under every candidate c1
(pass 2 at its default tile), then pass 2 under every candidate cr (pass 1 at
the chosen c1).  The times come from the plan's per-stage HIP events (summed
over a bucket's launches): the median of REPS executes after a warm-up.  A candidate replaces the default rule only when it
is at least MARGIN faster.  The bucket's cost is the time of both passes at
its tiles per grid cell; plans then give each curve the grid of least
cells x cost.  The table keeps every timing; build.py compiles the tiles and
costs into libgls.so.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

SIGMA = 1.35
M = 5
CELLS = 1.5e8        # grid cells per grid of a timed batch
POINTS_MIN, POINTS_MAX = 50, 5000   # points per curve (log-uniform, as the survey)
MIN_CELLS_PER_POINT = 20
POINT_LEVELS = 12
REPS = 5
MARGIN = 0.03        # a tile must beat the default rule by this fraction
POWER_DIFF_MAX = 1e-4   # powers of different tiles: FP32 atomics order only


def _use_sweep_library():
    lib = os.path.join(HERE, "libgls_sweep.so")
    if not os.path.exists(lib):
        raise SystemExit(f"{lib} not found: build it with `python -m cuvarbase_amd.build --sweep`")
    from . import _lib

    if _lib._lib is not None:
        raise RuntimeError("libgls is already loaded; run the autotuner in its own process")
    _lib.load(lib)


MIN_GRID, MAX_GRID = 1 << 10, 1 << 24


def all_buckets():
    """{(n1, p2): nf} for every bucket of the catalogue, with a frequency
    count whose grid need sigma * nf fits it (nf = n / sigma)."""
    from .build import ROW_LOG2, pass1_sizes

    out = {}
    for n1 in pass1_sizes():
        for p2 in ROW_LOG2:
            n = n1 << p2
            if MIN_GRID <= n <= MAX_GRID:
                out[(n1, p2)] = int(n / SIGMA)
    return out


def candidates(n1, p2):
    from .build import TILE_STEPS, pass1_tile_ok, pass2_tile_ok

    c1s = [c for c in TILE_STEPS if pass1_tile_ok(n1, c) and 2 * M + c <= (1 << p2)]
    crs = [c for c in TILE_STEPS if pass2_tile_ok(p2, c)]
    return c1s, crs


class BucketBench:
    """A synthetic batch of one size bucket on the device."""

    def __init__(self, n1, p2, nf, seed=0):
        from . import _lib
        from ._device import DeviceArray

        self.n1, self.p2, self.nf = n1, p2, nf
        n = n1 << p2
        self.ncurves = int(np.clip(round(CELLS / n), 4, 50_000))
        rng = np.random.default_rng(seed)
        T = 100.0
        df = 1.0 / (5.0 * T)
        f0 = 20.0 * df   # f (t_max - t_min) > 3 everywhere: no low-frequency direct sums
        # points per curve as in the survey workload: log-uniform over
        # [POINTS_MIN, POINTS_MAX] (at most one point per MIN_CELLS_PER_POINT
        # cells), drawn from POINT_LEVELS levels; each curve takes an evenly
        # spread subset of one sorted pool of times over [0, T)
        cap = max(POINTS_MIN, min(POINTS_MAX, n // MIN_CELLS_PER_POINT))
        levels = np.unique(np.geomspace(POINTS_MIN, cap, POINT_LEVELS).astype(int))
        pick = rng.integers(0, len(levels), size=self.ncurves)
        pool = np.sort(rng.uniform(0.0, T, int(levels[-1])))
        idx = {int(v): np.linspace(0, len(pool) - 1, int(v)).astype(np.int64) for v in levels}
        counts = levels[pick]
        t = np.concatenate([pool[idx[int(v)]] for v in counts])
        y = np.sin(2.0 * np.pi * t / 0.7) + rng.normal(0.0, 0.3, t.size)
        dy = 0.3 + 0.2 * rng.uniform(size=t.size)
        offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
        self.npts = float(np.mean(counts))
        self.curves = [_lib.gls_curve(obs_offset=int(o), n_obs=int(c), nf=nf, f0=f0, df=df,
                                      out_offset=i * nf, ybar=0.0, yy=1.0)
                       for i, (o, c) in enumerate(zip(offs, counts))]
        self.d_t = DeviceArray.to_device(t)
        self.d_a = DeviceArray.to_device(y)
        self.d_b = DeviceArray.to_device(dy)
        self.d_power = DeviceArray(self.ncurves * nf, np.float32)
        self.d_ws = None

    def time(self, c1, cr, reps=REPS):
        """(median pass-1 ms, median pass-2 ms, the first and the last
        curve's powers) with tiles (c1, cr)."""
        from . import _lib
        from ._device import DeviceArray, Stream, _handle
        from .engine import Plan, set_grid_cost, set_tiles

        # the tiles asked for, and this bucket made the cheapest grid for nf
        set_tiles(self.n1, self.p2, c1, cr)
        set_grid_cost(self.n1, self.p2, 1e-6)
        try:
            # spreading as the batched path's default picks it per curve
            plan = Plan.lomb(self.curves, SIGMA, M, _lib.MODE_FLOATING_MEAN, 0.0, 0.0,
                             _lib.INPUT_RAW, _lib.METHOD_NFFT, 0, _lib.WINDOW_ES, 0.0,
                             spread=_lib.SPREAD_AUTO)
        finally:
            set_tiles(self.n1, self.p2, 0, 0)
            set_grid_cost(self.n1, self.p2, 0.0)
        try:
            got = plan.tiles(0)
            if got != (c1, cr):
                raise RuntimeError(f"plan took tiles {got}, asked for {(c1, cr)}")
            geo = plan.geometry(0)
            if (geo[1], geo[2]) != (self.n1, 1 << self.p2):
                raise RuntimeError(f"plan took grid {geo[1]} x {geo[2]}, asked for "
                                   f"{self.n1} x {1 << self.p2}")
            need = plan.workspace_bytes
            if self.d_ws is None or self.d_ws.nbytes < need:
                self.d_ws = None
                self.d_ws = DeviceArray(need, np.uint8)
            plan.enable_timing(True)
            s = Stream()
            p1, p2 = [], []
            for r in range(reps + 1):
                _lib.call("gls_lomb_execute", plan.handle, self.d_t.ptr, self.d_a.ptr,
                          self.d_b.ptr, self.d_power.ptr, self.d_ws.ptr, _handle(s))
                s.synchronize()
                st = plan.stage_times()
                if r:
                    p1.append(sum(ms for name, _, ms, _ in st if name == "pass1"))
                    p2.append(sum(ms for name, _, ms, _ in st if name == "pass2"))
            # the first and the last curve: a launch that left curves out
            # (the end of its grid) shows in the last one
            ends = np.empty(2 * self.nf, dtype=np.float32)
            for k, c in enumerate((0, self.ncurves - 1)):
                _lib.call("gls_memcpy_d2h_async", ends[k * self.nf:].ctypes.data,
                          self.d_power.ptr + 4 * c * self.nf, C.c_size_t(4 * self.nf),
                          _handle(s))
            s.synchronize()
            return float(np.median(p1)), float(np.median(p2)), ends
        finally:
            plan.destroy()

    def release(self):
        self.d_t = self.d_a = self.d_b = self.d_power = self.d_ws = None


def pick(times, default):
    """The fastest tile, unless it beats the default by less than MARGIN."""
    best = min(times, key=times.get)
    if default in times and times[best] > (1.0 - MARGIN) * times[default]:
        return default
    return best


def tune_bucket(n1, p2, nf, reps=REPS, log=print):
    from .build import default_pass1_cols, default_pass2_rows

    c1s, crs = candidates(n1, p2)
    d1, d2 = default_pass1_cols(n1), default_pass2_rows(p2)
    bb = BucketBench(n1, p2, nf)
    ref, dmax = None, 0.0
    t1, t2 = {}, {}

    def run(c1_, cr_):
        # every tile computes the same periodogram (up to the order of the
        # FP32 spreading atomics): the first and the last curve's powers are
        # compared
        nonlocal ref, dmax
        a, b, first = bb.time(c1_, cr_, reps)
        if ref is None:
            ref = first
        dmax = max(dmax, float(np.max(np.abs(first - ref))))
        return a, b

    try:
        for c in c1s:
            t1[c] = run(c, d2)[0]
        c1 = pick(t1, d1)
        for c in crs:
            t2[c] = run(c1, c)[1]
        cr = pick(t2, d2)
    finally:
        bb.release()
    cells = bb.ncurves * 3 * (n1 << p2)
    rec = {"n1": n1, "p2": p2, "nf": nf, "curves": bb.ncurves, "points": bb.npts,
           "c1": c1, "cr": cr, "default": [d1, d2],
           "ps_per_cell": round((t1[c1] + t2[cr]) * 1e9 / cells, 4),
           "pass1_ms": {str(k): round(v, 4) for k, v in t1.items()},
           "pass2_ms": {str(k): round(v, 4) for k, v in t2.items()},
           "max_power_diff_between_tiles": dmax}
    log(f"bucket n1={n1} p2={p2} ({bb.ncurves} curves x {bb.npts:.0f} points, nf={nf}): "
        f"c1={c1} {rec['pass1_ms']}  cr={cr} {rec['pass2_ms']}  "
        f"{rec['ps_per_cell']} ps/cell", flush=True)
    if dmax > POWER_DIFF_MAX:
        log(f"WARNING: bucket n1={n1} p2={p2}: tiles disagree by {dmax:.3g} in the first or "
            "last curve's powers; its timings are not trusted", flush=True)
    return rec


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(HERE, "tiles_gfx950.json"))
    ap.add_argument("--reps", type=int, default=REPS)
    ap.add_argument("--max-buckets", type=int, default=0, help="tune only the first N (0: all)")
    args = ap.parse_args(argv)
    _use_sweep_library()
    from . import _lib

    _lib.init(0)
    t0 = time.time()
    buckets = all_buckets()
    print(f"{len(buckets)} size buckets", flush=True)
    keys = sorted(buckets)
    if args.max_buckets:
        keys = keys[:args.max_buckets]
    rows = [tune_bucket(n1, p2, buckets[(n1, p2)], args.reps) for n1, p2 in keys]
    doc = {"arch": "gfx950", "sigma": SIGMA, "m": M, "window": "exponential of semicircle",
           "cells_per_grid": CELLS, "points_per_curve": [POINTS_MIN, POINTS_MAX],
           "spread": "per curve (fixed point at one point per 80 cells or more, "
                     "on columns of 64 cells or more)",
           "min_cells_per_point": MIN_CELLS_PER_POINT, "reps": args.reps,
           "margin": MARGIN, "seconds": round(time.time() - t0, 1), "buckets": rows}
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    changed = sum(1 for r in rows if [r["c1"], r["cr"]] != r["default"])
    print(f"wrote {args.out}: {len(rows)} buckets, {changed} differ from the default rules",
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
