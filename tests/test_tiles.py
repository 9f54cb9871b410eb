"""Tile choice of the NFFT passes: the measured table, the generated header the
library is compiled from, and the host-side selection (gls_tile_get/set, which
need no device)."""
import json
import os

import pytest

from cuvarbase_amd import build

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_generated_header_is_current():
    """csrc/tiles_gen.hpp is what build.py makes from the committed table."""
    path = os.path.join(ROOT, "cuvarbase_amd", "csrc", "tiles_gen.hpp")
    assert open(path).read() == build.tile_header()


def test_header_lists_every_default_and_table_tile():
    text = build.tile_header(table=[(4096, 7, 2, 16, 4.5)])
    p1 = text.split("GLS_PASS1_TILES(X)")[1].split("#define")[0]
    p2 = text.split("GLS_PASS2_TILES(X)")[1].split("#define")[0]
    for n1 in build.pass1_sizes():
        assert f"X({n1}, {build.default_pass1_cols(n1)})" in p1
    for p in build.ROW_LOG2:
        assert f"X({p}, {build.default_pass2_rows(p)})" in p2
    assert "X(4096, 2)" in p1 and "X(7, 16)" in p2
    assert "X(4096, 7, 2, 16, 4.5)" in text.split("GLS_TILE_TABLE(X)")[1]
    # the model scale of a one-bucket table: its cost over its model units
    assert f"GLS_MODEL_PS_PER_UNIT {round(4.5 / build.model_units(4096, 7), 4)!r}" in text
    # no duplicate kernel (duplicate pairs would be duplicate dispatch cases)
    assert p1.count("X(4096, 1)") == 1


def test_sweep_header_has_every_candidate():
    text = build.tile_header(sweep=True)
    assert "GLS_TILE_TABLE(X)\n" in text   # no table in the sweep library
    n_p1 = sum(1 for n1 in build.pass1_sizes() for c in build.TILE_STEPS
               if build.pass1_tile_ok(n1, c))
    n_p2 = sum(1 for p in build.ROW_LOG2 for c in build.TILE_STEPS if build.pass2_tile_ok(p, c))
    assert text.count("X(") == n_p1 + n_p2
    # the candidate rules: n1 c1 <= 8192, 256 <= 2^p2 cr <= 4096
    assert build.pass1_tile_ok(4096, 2) and not build.pass1_tile_ok(4096, 4)
    assert build.pass2_tile_ok(6, 4) and not build.pass2_tile_ok(6, 2)
    assert build.pass2_tile_ok(12, 1) and not build.pass2_tile_ok(12, 2)


def test_radix_stages_follow_the_fft_radix_order():
    # 4096 = 16^3, 2304 = 16 16 9, 405 = 9 9 5, 27 = 9 3, 128 = 16 8, 20 = 4 5
    assert [build.radix_stages(v) for v in (4096, 2304, 405, 27, 128, 20, 1)] == \
        [3, 3, 3, 2, 2, 2, 0]
    assert build.model_units(4096, 7) == pytest.approx(1.0 + 0.12 * 5)


def test_default_rules_match_the_survey_tuned_split():
    # tiles of 4096 cells for pass 1 (at most 16 columns), pass-2 tiles of
    # 1024 row outputs up to rows of 256, 2048 beyond
    assert [build.default_pass1_cols(n) for n in (16, 256, 384, 1280, 2304, 4096)] == \
        [16, 16, 8, 2, 1, 1]
    assert [build.default_pass2_rows(p) for p in range(6, 13)] == [16, 8, 4, 4, 2, 1, 1]


def _table():
    if not os.path.exists(build.TILE_TABLE):
        pytest.skip("no measured tile table")
    return json.load(open(build.TILE_TABLE))


def test_table_choices_are_measured_minima():
    doc = _table()
    assert doc["buckets"], "empty table"
    margin = doc["margin"]
    for b in doc["buckets"]:
        t1 = {int(k): v for k, v in b["pass1_ms"].items()}
        t2 = {int(k): v for k, v in b["pass2_ms"].items()}
        d1, d2 = b["default"]
        assert build.pass1_tile_ok(b["n1"], b["c1"]) and build.pass2_tile_ok(b["p2"], b["cr"])
        for t, chosen, default in ((t1, b["c1"], d1), (t2, b["cr"], d2)):
            assert chosen in t
            best = min(t.values())
            if chosen == default:
                assert best >= (1 - margin) * t[default] - 1e-9
            else:
                assert t[chosen] == best and best < (1 - margin) * t[default]
        # every candidate tile computed the same powers (FP32 atomics order aside)
        assert b["max_power_diff_between_tiles"] < 1e-4
        # the bucket's cost: both passes at its tiles per grid cell
        cells = b["curves"] * 3 * (b["n1"] << b["p2"])
        assert b["ps_per_cell"] == pytest.approx((t1[b["c1"]] + t2[b["cr"]]) * 1e9 / cells,
                                                 rel=1e-3)
    # the whole catalogue is measured (1024 <= n <= 2^24)
    measured = {(b["n1"], b["p2"]) for b in doc["buckets"]}
    assert measured == {(n1, p) for n1 in build.pass1_sizes() for p in build.ROW_LOG2
                        if 1024 <= (n1 << p) <= 1 << 24}


def test_library_tiles_follow_table_and_rules():
    from cuvarbase_amd import _lib, engine

    try:
        _lib.load()
    except _lib.GLSLibraryError:
        pytest.skip("libgls.so not built")
    table = {(n1, p2): (c1, cr) for n1, p2, c1, cr, _ in build.load_tile_table()}
    for n1 in build.pass1_sizes():
        for p2 in build.ROW_LOG2:
            want = table.get((n1, p2), (build.default_pass1_cols(n1),
                                        build.default_pass2_rows(p2)))
            assert engine.get_tiles(n1, p2) == want, (n1, p2)


def test_tile_override_accepts_built_and_rejects_missing_kernels():
    from cuvarbase_amd import _lib, engine

    try:
        _lib.load()
    except _lib.GLSLibraryError:
        pytest.skip("libgls.so not built")
    base = engine.get_tiles(4096, 7)
    with pytest.raises(_lib.GLSError):
        engine.set_tiles(4096, 7, 16, 0)       # 65536-cell tile: not a candidate
    with pytest.raises(_lib.GLSError):
        engine.set_tiles(4095, 7, 1, 0)        # not a catalogue size
    with pytest.raises(_lib.GLSError):
        engine.set_tiles(4096, 13, 1, 0)       # rows longer than 4096
    assert engine.get_tiles(4096, 7) == base
    # a built pass-2 tile of another bucket with the same row length
    engine.set_tiles(4096, 7, 0, build.default_pass2_rows(7))
    try:
        assert engine.get_tiles(4096, 7) == (base[0], build.default_pass2_rows(7))
    finally:
        engine.set_tiles(4096, 7, 0, 0)
    assert engine.get_tiles(4096, 7) == base


def test_library_grid_costs_follow_table_and_model():
    from cuvarbase_amd import _lib, engine

    try:
        _lib.load()
    except _lib.GLSLibraryError:
        pytest.skip("libgls.so not built")
    rows = build.load_tile_table()
    table = {(n1, p2): ps for n1, p2, _, _, ps in rows}
    scale = build.model_scale(rows)
    for n1 in build.pass1_sizes():
        for p2 in build.ROW_LOG2:
            c, src = engine.get_grid_cost(n1, p2)
            if (n1, p2) in table:
                assert src == 1 and c == pytest.approx(table[(n1, p2)])
            else:
                assert src == 0 and c == pytest.approx(scale * build.model_units(n1, p2))
    base = engine.get_grid_cost(4096, 7)
    engine.set_grid_cost(4096, 7, 0.5)
    try:
        assert engine.get_grid_cost(4096, 7) == (0.5, 2)
    finally:
        engine.set_grid_cost(4096, 7, 0.0)
    assert engine.get_grid_cost(4096, 7) == base
    with pytest.raises(_lib.GLSError):
        engine.set_grid_cost(4095, 7, 1.0)
