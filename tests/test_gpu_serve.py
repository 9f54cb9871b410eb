"""Continuous-batching server on the GPU.

The server spreads in fixed point by default, so a request's periodogram is
bitwise the same served in a mix and computed alone with the same setting
(integer accumulation is exact and order-free).  MIX_TOL = 2e-5 is the bar
where FP32-atomic spreading (GLS_SPREAD_FLOAT, the batched engine's default)
is involved: the order of its additions changes with the batch and the
deconvolution amplifies that round-off at the band edges.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIX_TOL = 2e-5


def _stream(n=40, rate=2000.0, seed=41):
    from cuvarbase_amd import synthetic

    return synthetic.stream_requests(rate, nreq=n, seed=seed)


def test_each_result_same_alone_and_in_the_mix():
    from cuvarbase_amd import engine
    from cuvarbase_amd.serve import GLSServer

    arrivals, data, freqs = _stream()
    srv = GLSServer(hbm_budget=8e9, slots=3, max_batch_evals=2e7)
    st = srv.serve_stream(arrivals, data, freqs)
    assert st["batches"] > 1
    worst = 0.0
    for r, d, f in zip(st["requests_obj"], data, freqs):
        assert r.done and r.latency > 0
        alone = engine.LombBatch([d], [f], sigma=srv.sigma, m=srv.m, window=srv.window,
                                 window_beta=srv.window_beta, spread=srv.spread).run()[0]
        worst = max(worst, float(np.max(np.abs(r.power.astype(float) - alone))))
        assert np.array_equal(r.power, alone)
    print(f"served {st['requests']} requests in {st['batches']} batches, p50 {st['p50_ms']:.1f} "
          f"ms, p99 {st['p99_ms']:.1f} ms; alone vs mix max|dP| = {worst:.2e}")


def test_admission_waits_for_memory():
    from cuvarbase_amd.engine import LombBatch
    from cuvarbase_amd.serve import GLSServer

    arrivals, data, freqs = _stream(n=12, seed=42)
    # budget for about two of the largest requests at a time
    from cuvarbase_amd.tuning import choose_nfft_params

    sigma, m = choose_nfft_params(1e-4)
    sizes = []
    for d, f in zip(data, freqs):
        b = LombBatch([d], [f], sigma=sigma, m=m)
        sizes.append(b.plan.workspace_bytes + 24 * len(d[0]) + 4 * len(f) + 3 * 256)
    srv = GLSServer(hbm_budget=2.2 * max(sizes), slots=4, max_batch_evals=1e9, sigma=sigma,
                    m=m)
    for (t, y, dy), f in zip(data, freqs):
        srv.submit(t, y, dy, f)
    peak = 0
    while srv.busy:
        srv.poll()
        peak = max(peak, srv.arena.used)
    assert peak <= srv.arena.nbytes
    assert len(srv.completed) == len(data)
    assert srv.batches_launched >= 2


def test_request_larger_than_budget_raises():
    from cuvarbase_amd.serve import GLSServer

    _, data, freqs = _stream(n=1, seed=43)
    srv = GLSServer(hbm_budget=4096, slots=1)
    t, y, dy = data[0]
    srv.submit(t, y, dy, freqs[0])
    with pytest.raises(MemoryError):
        srv.drain()


def test_serve_records_mixed_classes():
    """Linear full-periodogram, linear top-k and custom-frequency requests in
    one record stream: each result equals the same request computed alone."""
    import io

    from cuvarbase_amd import engine, wire
    from cuvarbase_amd.lombscargle import fap_baluev
    from cuvarbase_amd.serve import GLSServer

    _, data, freqs = _stream(n=9, seed=44)
    recs = []
    rng = np.random.default_rng(5)
    for i, ((t, y, dy), f) in enumerate(zip(data, freqs)):
        if i % 3 == 0:
            recs.append(wire.RequestRecord(100 + i, t, y, dy, f0=f[0], df=f[1] - f[0],
                                           nf=len(f)))
        elif i % 3 == 1:
            recs.append(wire.RequestRecord(100 + i, t, y, dy, f0=f[0], df=f[1] - f[0],
                                           nf=len(f), k=5))
        else:
            recs.append(wire.RequestRecord(100 + i, t, y, dy,
                                           freqs=np.sort(rng.uniform(0.01, 5.0, 2000)), k=0))
    src = io.BytesIO(b"".join(wire.encode_request(r) for r in recs))
    dst = io.BytesIO()
    srv = GLSServer(hbm_budget=4e9, slots=2, max_batch_evals=5e6)
    assert srv.serve_records(src, dst) == len(recs)
    dst.seek(0)
    res = {r.id: r for r in wire.iter_results(dst)}
    assert sorted(res) == [r.id for r in recs]
    for r in recs:
        out = res[r.id]
        assert out.status == 0 and out.latency > 0
        if r.freqs is not None:
            alone = engine.LombBatch([(r.t, r.y, r.dy)], [r.freqs], custom_freqs=True).run()[0]
            np.testing.assert_allclose(out.power, alone, atol=1e-6)
            continue
        alone = engine.LombBatch([(r.t, r.y, r.dy)], [r.frequencies()], sigma=srv.sigma,
                                 m=srv.m, window=srv.window,
                                 window_beta=srv.window_beta).run()[0]
        if r.k == 0:
            assert np.max(np.abs(out.power - alone)) < MIX_TOL
        else:
            # the served best peak is the best alone (up to MIX_TOL ties)
            assert alone[out.top_index[0]] >= alone.max() - MIX_TOL
            assert np.all(np.diff(out.top_power) <= 0)
            np.testing.assert_allclose(out.top_power, alone[out.top_index], atol=MIX_TOL)
            np.testing.assert_allclose(out.top_freq, r.frequencies()[out.top_index])
            fap = fap_baluev(r.t, r.dy, out.top_power.astype(np.float64), r.frequencies()[-1])
            np.testing.assert_allclose(out.fap, fap, rtol=1e-9, atol=1e-12)


def test_zero_copy_results_and_release():
    """Zero-copy mode: results are views into the page-locked result arena,
    equal to the copy mode's, and release() returns the memory."""
    from cuvarbase_amd.serve import GLSServer

    arrivals, data, freqs = _stream(n=10, seed=45)
    a = GLSServer(hbm_budget=4e9, slots=2, max_batch_evals=3e6)
    b = GLSServer(hbm_budget=4e9, slots=2, max_batch_evals=3e6, zero_copy=True)
    sa = a.serve_stream(arrivals * 0, data, freqs)
    sb = b.serve_stream(arrivals * 0, data, freqs)
    assert b.host_arena.used > 0
    for ra, rb in zip(sa["requests_obj"], sb["requests_obj"]):
        assert rb.power.base is not None            # a view, not a copy
        assert np.max(np.abs(ra.power - rb.power)) < MIX_TOL
        rb.release()
        assert rb.power is None
    assert b.host_arena.used == 0


def test_graph_served_top5_matches_eager_and_oracle():
    """The short-curve workload through the server on one shared grid, k = 5:
    batches replayed from a captured execute (rebound plan, descriptors
    uploaded per batch) against eager batches and against the FP64 oracle's
    top 5 (ties within 1e-4 admitted), as in config 5."""
    from oracle import gls as og

    from cuvarbase_amd import synthetic
    from cuvarbase_amd.serve import GLSServer

    data, freqs = synthetic.short(ncurves=24)
    grid = freqs[0]

    def served(graphs):
        srv = GLSServer(hbm_budget=30e9, slots=1, max_batch_curves=8, graph_batches=graphs)
        reqs = [srv.submit(*d, grid, k=5) for d in data]
        srv.drain()
        return srv, reqs

    sg, rg = served(True)
    se, re_ = served(False)
    # three batches of one shape: eager, captured, replayed
    assert sg.graph_captures == 1 and sg.graph_replays == 2
    assert se.graph_captures == 0 and se.graph_replays == 0
    for i, (a, b, d) in enumerate(zip(rg, re_, data)):
        np.testing.assert_array_equal(a.top_power, b.top_power)
        np.testing.assert_array_equal(a.top_index, b.top_index)
        Po = og.gls_direct(*d, grid)
        order = np.argsort(-Po, kind="stable")[:5]
        for q in range(5):
            j = int(a.top_index[q])
            assert abs(float(a.top_power[q]) - Po[j]) <= 1e-4, (i, q)
            assert j == int(order[q]) or abs(Po[j] - Po[order[q]]) <= 1e-4, (i, q, j, order)
            jb = int(b.top_index[q])
            assert j == jb or abs(Po[j] - Po[jb]) <= 1e-4
        assert np.all((a.fap >= 0) & (a.fap <= 1))


def test_fixed_point_spreading_is_bitwise_reproducible():
    """GLS_SPREAD_FIXED: every request's periodogram served in a mix is
    bitwise equal to the same curve computed alone (and to a second run),
    curves above 16,384 points included, and meets 1e-4 against FP64."""
    from oracle import gls as og

    from cuvarbase_amd import _lib, engine
    from cuvarbase_amd.serve import GLSServer

    arrivals, data, freqs = _stream()
    assert sum(len(d[0]) > 16384 for d in data) >= 3
    srv = GLSServer(hbm_budget=8e9, slots=3, max_batch_evals=2e7, spread=_lib.SPREAD_FIXED)
    st = srv.serve_stream(arrivals, data, freqs)
    assert st["batches"] > 1
    cfg = dict(sigma=srv.sigma, m=srv.m, window=srv.window, window_beta=srv.window_beta)
    rng = np.random.default_rng(5)
    again = engine.LombBatch(data, freqs, spread=_lib.SPREAD_FIXED, **cfg).run()
    worst = 0.0
    for i, (r, d, f) in enumerate(zip(st["requests_obj"], data, freqs)):
        alone = engine.LombBatch([d], [f], spread=_lib.SPREAD_FIXED, **cfg).run()[0]
        assert np.array_equal(r.power, alone), i
        assert np.array_equal(again[i], alone), i
        if i % 4 == 0:
            sel = np.unique(np.concatenate([np.arange(20), [int(np.argmax(alone))],
                                            rng.choice(len(f), 40, replace=False)]))
            worst = max(worst, float(np.max(np.abs(alone[sel] - og.gls_direct(*d, f[sel])))))
    print(f"fixed-point spreading: served == alone bitwise for {len(data)} requests; "
          f"max|dP| vs FP64 {worst:.2e}")
    assert worst < 1e-4


def test_float_spreading_alone_vs_mix_within_tolerance():
    """GLS_SPREAD_FLOAT through the server: results of the mix equal the
    curve alone within MIX_TOL (FP32 atomic order)."""
    from cuvarbase_amd import _lib, engine
    from cuvarbase_amd.serve import GLSServer

    arrivals, data, freqs = _stream(n=24, seed=43)
    srv = GLSServer(hbm_budget=8e9, slots=3, max_batch_evals=2e7, spread=_lib.SPREAD_FLOAT)
    st = srv.serve_stream(arrivals, data, freqs)
    worst = 0.0
    for r, d, f in zip(st["requests_obj"], data, freqs):
        alone = engine.LombBatch([d], [f], sigma=srv.sigma, m=srv.m, window=srv.window,
                                 window_beta=srv.window_beta, spread=_lib.SPREAD_FLOAT).run()[0]
        worst = max(worst, float(np.max(np.abs(r.power.astype(float) - alone))))
    print(f"float spreading: alone vs mix max|dP| = {worst:.2e}")
    assert worst < MIX_TOL
