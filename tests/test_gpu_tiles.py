"""The tile choice of the NFFT passes changes the work split, not the result.

For buckets where the measured table picked a tile other than the default
rule, the library has kernels for both; the same batch run with either must
give the same periodogram: bitwise when only the pass-2 tile differs (rows are
transformed independently, with the same arithmetic, whatever the tile), and
within the FP32 rounding of the spreading when the pass-1 tile differs (the
fixed-point spreading picks its exponents per tile row)."""
import numpy as np
import pytest

from cuvarbase_amd import _lib, build, engine

pytestmark = pytest.mark.gpu

SAME_TOL = 2e-5


def _differing_buckets():
    nf_of = {}
    import json
    import os

    if os.path.exists(build.TILE_TABLE):
        for b in json.load(open(build.TILE_TABLE))["buckets"]:
            nf_of[(b["n1"], b["p2"])] = b["nf"]
    out = []
    for n1, p2, c1, cr, _ in build.load_tile_table():
        d1, d2 = build.default_pass1_cols(n1), build.default_pass2_rows(p2)
        if (c1, cr) != (d1, d2):
            out.append((n1, p2, c1, cr, d1, d2, nf_of[(n1, p2)]))
    return out


def _batch(nf, ncurves=3, npts=300, seed=5):
    rng = np.random.default_rng(seed)
    data, freqs = [], []
    df = 1.0 / (5 * 200.0)
    fr = df * (25 + np.arange(nf))
    for _ in range(ncurves):
        t = np.sort(rng.uniform(0, 200.0, npts))
        y = np.sin(2 * np.pi * t / 1.7) + rng.normal(0, 0.4, npts)
        dy = 0.3 + 0.3 * rng.uniform(size=npts)
        data.append((t, y, dy))
        freqs.append(fr)
    return data, freqs


def _run(data, freqs, n1, p2):
    # the bucket made the cheapest grid for these curves while planning
    engine.set_grid_cost(n1, p2, 1e-6)
    try:
        b = engine.LombBatch(data, freqs, sigma=1.35, m=5, window=_lib.WINDOW_ES,
                             spread=_lib.SPREAD_FIXED)
    finally:
        engine.set_grid_cost(n1, p2, 0.0)
    geo = b.plan.geometry(0)
    assert (geo[1], geo[2]) == (n1, 1 << p2), "batch is not in the bucket"
    return b.plan.tiles(0), np.concatenate(b.run())


def test_tile_choice_does_not_change_powers():
    cases = _differing_buckets()
    if not cases:
        pytest.skip("the measured table keeps the default tiles everywhere")
    # the three largest such buckets (most rows and columns per curve)
    cases.sort(key=lambda c: -(c[0] << c[1]))
    for n1, p2, c1, cr, d1, d2, nf in cases[:3]:
        print(f"bucket n1={n1} p2={p2}: table ({c1}, {cr}) vs default ({d1}, {d2})", flush=True)
        data, freqs = _batch(nf)
        tiles_t, p_table = _run(data, freqs, n1, p2)
        assert tiles_t == (c1, cr)
        # the other pass-2 tile alone: bitwise
        if cr != d2:
            engine.set_tiles(n1, p2, c1, d2)
            try:
                tiles_d, p_d = _run(data, freqs, n1, p2)
            finally:
                engine.set_tiles(n1, p2, 0, 0)
            assert tiles_d == (c1, d2)
            assert np.array_equal(p_table, p_d)
        # the other pass-1 tile: same powers to FP32 rounding
        if c1 != d1:
            engine.set_tiles(n1, p2, d1, cr)
            try:
                tiles_d, p_d = _run(data, freqs, n1, p2)
            finally:
                engine.set_tiles(n1, p2, 0, 0)
            assert tiles_d == (d1, cr)
            assert np.max(np.abs(p_table - p_d)) < SAME_TOL
