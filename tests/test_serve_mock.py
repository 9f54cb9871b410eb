"""Host logic of the continuous-batching server against a simulated device.

The libgls entry points the server calls are replaced by a host stand-in in
which every stream is a queue that runs only when an event on it is queried
(at random) or synchronised, so work completes late and out of order as on a
GPU: a server that reused a staging buffer, a result block or an arena block
too early, mixed up offsets between requests, or retired a batch before its
copies ran would produce wrong periodograms here.  The simulated execute
computes each curve with the FP64 oracle's direct sums; top-k uses a stable
argsort.  No GPU is involved.
"""
import ctypes as C
import io
import random

import numpy as np
import pytest

from oracle import gls as og


class FakeDevice:
    def __init__(self, seed=1):
        self.rng = random.Random(seed)
        self.keep, self.plans, self.streams, self.events = {}, {}, {}, {}
        self.graphs, self.ws_meta = {}, {}
        self.next_plan = 1000
        self.next_graph = 5000

    def _enq(self, s, fn):
        self.streams.setdefault(s, []).append(fn)

    def _run(self, s):
        q = self.streams.get(s, [])
        while q:
            q.pop(0)()

    def _arr(self, ctype, addr, n):
        return np.ctypeslib.as_array((ctype * n).from_address(addr)) if n else np.zeros(0)

    def call(self, name, *args):
        a = [x.value if isinstance(x, (C.c_int32, C.c_size_t, C.c_void_p, C.c_double)) else x
             for x in args]
        if name == "gls_init":
            return 0
        if name == "gls_stream_create":
            h = 0x10000 + len(self.streams) + 1
            self.streams[h] = []
            args[0]._obj.value = h
            return 0
        if name == "gls_event_create":
            h = 0x20000 + len(self.events) + 1
            self.events[h] = [True, None]
            args[0]._obj.value = h
            return 0
        if name in ("gls_stream_destroy", "gls_event_destroy", "gls_free", "gls_host_free"):
            return 0
        if name == "gls_event_record":
            ev, s = a
            self.events[ev] = [False, s]
            self._enq(s, lambda: self.events[ev].__setitem__(0, True))
            return 0
        if name == "gls_event_query":
            ev = a[0]
            if not self.events[ev][0] and self.rng.random() < 0.3:
                self._run(self.events[ev][1])
            args[1]._obj.value = int(self.events[ev][0])
            return 0
        if name == "gls_stream_synchronize":
            self._run(a[0])
            return 0
        if name in ("gls_device_synchronize", "gls_event_synchronize"):
            for s in list(self.streams):
                self._run(s)
            return 0
        if name in ("gls_malloc", "gls_host_alloc"):
            buf = C.create_string_buffer(max(int(a[1]), 1))
            self.keep[C.addressof(buf)] = buf
            args[0]._obj.value = C.addressof(buf)
            return 0
        if name.startswith("gls_memcpy"):
            dst, src, n, s = a
            self._enq(s, lambda: C.memmove(dst, src, n))
            return 0
        if name == "gls_lomb_plan_create":
            h, arr, n, prm = args
            self.next_plan += 1
            self.plans[self.next_plan] = dict(
                curves=[(c.obs_offset, c.n_obs, c.nf, c.f0, c.df, c.out_offset)
                        for c in (arr[i] for i in range(n.value))],
                mode=prm._obj.mode, alive=True)
            h._obj.value = self.next_plan
            return 0
        if name == "gls_plan_workspace_bytes":
            args[1]._obj.value = 4096
            return 0
        if name == "gls_plan_destroy":
            self.plans[a[0]]["alive"] = False
            return 0
        if name in ("gls_lomb_execute", "gls_lomb_execute_custom"):
            if name == "gls_lomb_execute":
                h, t, y, dy, pw, ws, s = a
                fr = None
            else:
                h, t, y, dy, fr, pw, _f64, ws, s = a
            # an execute uploads the plan's descriptors (as they are at the
            # call) into the workspace, then computes from them
            meta = list(self.plans[h]["curves"])
            mode = self.plans[h]["mode"]

            def run():
                assert self.plans[h]["alive"], "plan destroyed before its execute ran"
                C.memset(ws, 0x7F, 4096)   # a shared workspace would be clobbered
                self.ws_meta[ws] = meta
                self._compute(meta, mode, t, y, dy, fr, pw)
            self._enq(s, run)
            return 0
        if name == "gls_plan_graph_create":
            h, t, y, dy, pw, _f64, ws, s, out = args
            h, t, y, dy, pw, ws, s = (x.value if isinstance(x, C.c_void_p) else x
                                      for x in (h, t, y, dy, pw, ws, s))
            self._run(s)   # creation synchronises its stream
            self.ws_meta[ws] = list(self.plans[h]["curves"])
            self.next_graph += 1
            self.graphs[self.next_graph] = dict(plan=h, ptrs=(t, y, dy, pw), ws=ws, alive=True,
                                                mode=self.plans[h]["mode"])
            out._obj.value = self.next_graph
            return 0
        if name == "gls_plan_rebind":
            h, arr, n = args
            h = h.value if isinstance(h, C.c_void_p) else h
            P = self.plans[h]
            assert P["alive"] and n.value == len(P["curves"])
            new = [(c.obs_offset, c.n_obs, c.nf, c.f0, c.df, c.out_offset)
                   for c in (arr[i] for i in range(n.value))]
            assert [c[2:5] for c in new] == [c[2:5] for c in P["curves"]], "grids changed"
            P["curves"] = new
            return 0
        if name == "gls_graph_update":
            g, h, s = a
            G = self.graphs[g]
            assert G["alive"] and G["plan"] == h
            meta = list(self.plans[h]["curves"])
            self._enq(s, lambda: self.ws_meta.__setitem__(G["ws"], meta))
            return 0
        if name == "gls_graph_launch":
            g, s = a
            G = self.graphs[g]

            def replay():
                assert G["alive"], "graph destroyed before its replay ran"
                t, y, dy, pw = G["ptrs"]
                self._compute(self.ws_meta[G["ws"]], G["mode"], t, y, dy, None, pw)
            self._enq(s, replay)
            return 0
        if name == "gls_graph_destroy":
            self.graphs[a[0]]["alive"] = False
            return 0
        if name == "gls_topk":
            h, ws, pw, mask, k, tp, ti, s = a

            def run():
                for ci, (o, n, nf, f0, df, oo) in enumerate(self.ws_meta[ws]):
                    p = self._arr(C.c_float, pw + 4 * oo, nf).copy()
                    order = np.argsort(-p, kind="stable")[:k]
                    self._arr(C.c_float, tp + 4 * ci * k, k)[:] = p[order]
                    self._arr(C.c_int32, ti + 4 * ci * k, k)[:] = order
            self._enq(s, run)
            return 0
        raise AssertionError(f"unexpected libgls call {name}")

    def _compute(self, meta, mode, t, y, dy, fr, pw):
        for (o, n, nf, f0, df, oo) in meta:
            tt = self._arr(C.c_double, t + 8 * o, n).copy()
            yy = self._arr(C.c_double, y + 8 * o, n).copy()
            dd = self._arr(C.c_double, dy + 8 * o, n).copy()
            f = (f0 + df * np.arange(nf) if fr is None
                 else self._arr(C.c_double, fr + 8 * oo, nf).copy())
            self._arr(C.c_float, pw + 4 * oo, nf)[:] = og.gls_direct(tt, yy, dd, f, mode=mode)


@pytest.fixture
def serve(monkeypatch):
    from cuvarbase_amd import _lib, serve as srvmod

    dev = FakeDevice()
    monkeypatch.setattr(_lib, "call", dev.call)
    monkeypatch.setattr(_lib, "init", lambda device=0: None)
    monkeypatch.setattr(srvmod.time, "sleep", lambda s: None)
    monkeypatch.setattr(srvmod, "_test_device", dev, raising=False)
    yield srvmod
    # release the simulated streams, events and buffers while the stand-in is
    # still in place (their handles mean nothing to the real library)
    import gc
    gc.collect()


def _requests(n, seed=3):
    rng = np.random.default_rng(seed)
    data, freqs = [], []
    for i in range(n):
        m = int(rng.integers(20, 80))
        t = np.sort(rng.uniform(0, 20, m))
        data.append((t, np.cos(t * (2 + i)) + rng.normal(0, 0.3, m), np.full(m, 0.3)))
        freqs.append(0.05 + 0.01 * np.arange(int(rng.integers(200, 400))))
    return data, freqs


@pytest.mark.parametrize("zero_copy", [False, True])
def test_stream_results_match_oracle(serve, zero_copy):
    data, freqs = _requests(24)
    srv = serve.GLSServer(hbm_budget=1 << 22, slots=3, max_batch_evals=700,
                          zero_copy=zero_copy, result_buffer_bytes=1 << 16)
    st = srv.serve_stream(np.zeros(len(data)), data, freqs)
    assert st["batches"] > 3
    for r, d, f in zip(st["requests_obj"], data, freqs):
        np.testing.assert_allclose(r.power, og.gls_direct(*d, f), atol=1e-6)
        r.release()
    if zero_copy:
        assert srv.host_arena.used == 0


def test_records_mixed_classes_and_release(serve):
    from cuvarbase_amd import wire

    data, freqs = _requests(18, seed=4)
    rng = np.random.default_rng(9)
    recs = []
    for i, (d, f) in enumerate(zip(data, freqs)):
        if i % 3 == 2:
            recs.append(wire.RequestRecord(i, *d, freqs=np.sort(rng.uniform(0.05, 3, 150))))
        else:
            recs.append(wire.RequestRecord(i, *d, f0=f[0], df=f[1] - f[0], nf=len(f),
                                           k=3 if i % 3 == 1 else 0))
    srv = serve.GLSServer(hbm_budget=1 << 22, slots=2, max_batch_evals=700, zero_copy=True,
                          result_buffer_bytes=1 << 14)
    dst = io.BytesIO()
    src = io.BytesIO(b"".join(wire.encode_request(r) for r in recs))
    assert srv.serve_records(src, dst) == len(recs)
    assert srv.host_arena.used == 0
    dst.seek(0)
    out = {r.id: r for r in wire.iter_results(dst)}
    for rq in recs:
        ref = og.gls_direct(rq.t, rq.y, rq.dy, rq.frequencies())
        o = out[rq.id]
        if rq.k:
            order = np.argsort(-ref.astype(np.float32), kind="stable")[:3]
            np.testing.assert_array_equal(o.top_index, order)
            np.testing.assert_allclose(o.top_freq, rq.frequencies()[order])
            assert np.all((o.fap >= 0) & (o.fap <= 1))
        else:
            np.testing.assert_allclose(o.power, ref, atol=1e-6)


def test_unreleased_results_block_admission(serve):
    data, freqs = _requests(6, seed=5)
    srv = serve.GLSServer(hbm_budget=1 << 22, slots=1, max_batch_evals=300, zero_copy=True,
                          result_buffer_bytes=4 * 420)
    st = srv.serve_stream(np.zeros(1), data[:1], freqs[:1])
    assert srv.host_arena.used > 0          # the first result is still held
    t, y, dy = data[1]
    srv.submit(t, y, dy, freqs[1])
    with pytest.raises(MemoryError, match="release"):
        srv.drain()
    st["requests_obj"][0].release()
    assert srv.host_arena.used == 0


def test_batches_respect_point_cap(serve):
    data, freqs = _requests(20, seed=6)
    cap = 150
    srv = serve.GLSServer(hbm_budget=1 << 22, slots=2, max_batch_evals=10**6,
                          max_batch_points=cap)
    assert all(s.h_in.size >= 3 * cap for s in srv.slots)   # staging sized up front
    before = set(serve._test_device.plans)                  # the start-up warm-up plan
    st = srv.serve_stream(np.zeros(len(data)), data, freqs)
    plans = [p for h, p in serve._test_device.plans.items() if h not in before]
    assert len(plans) == st["batches"] > 1
    for p in plans:
        pts = sum(c[1] for c in p["curves"])
        assert pts <= cap or len(p["curves"]) == 1
    for r, d, f in zip(st["requests_obj"], data, freqs):
        np.testing.assert_allclose(r.power, og.gls_direct(*d, f), atol=1e-6)


def test_server_starts_clean_after_warm_up(serve):
    """The start-up warm-up request leaves no trace: no results, no batches
    counted, every arena byte returned."""
    srv = serve.GLSServer(hbm_budget=1 << 22, slots=2, max_batch_evals=10**4, zero_copy=True,
                          result_buffer_bytes=4 * 4096)
    assert srv.completed == [] and not srv.busy
    assert srv.batches_launched == 0 and srv.evals_done == 0
    assert srv.arena.used == 0 and srv.host_arena.used == 0


def test_tiny_budget_skips_warm_up(serve):
    """A budget too small for the warm-up request still builds a server; the
    first real request that does not fit raises as before."""
    srv = serve.GLSServer(hbm_budget=256, slots=1)
    assert srv.completed == [] and not srv.busy and srv.arena.used == 0
    data, freqs = _requests(1, seed=9)
    t, y, dy = data[0]
    srv.submit(t, y, dy, freqs[0])
    with pytest.raises(MemoryError, match="budget"):
        srv.drain()


def test_shared_grid_batches_replay_captured_graphs(serve):
    """Requests on one linear grid: the first batch of a shape runs eagerly,
    the second captures a graph on its slot, later batches of the shape
    rebind the plan and replay it (inputs at the captured row stride,
    descriptors uploaded before the replay); top-k and whole periodograms
    match the oracle, and a batch of another size is a new shape."""
    rng = np.random.default_rng(11)
    grid = 0.05 + 0.01 * np.arange(300)
    data = []
    for i in range(14):
        m = int(rng.integers(20, 90))
        t = np.sort(rng.uniform(0, 20, m))
        data.append((t, np.cos(t * (1.5 + 0.3 * i)) + rng.normal(0, 0.3, m), np.full(m, 0.3)))
    for k in (0, 4):
        srv = serve.GLSServer(hbm_budget=1 << 24, slots=1, max_batch_curves=4,
                              max_batch_points=400)
        reqs = [srv.submit(*d, grid, k=k) for d in data]
        srv.drain()
        assert srv.batches_launched == 4          # 4 + 4 + 4 + 2 curves
        # shape 4: eager, captured, replayed; shape 2: seen once, eager
        assert srv.graph_captures == 1 and srv.graph_replays == 2
        for r, d in zip(reqs, data):
            ref = og.gls_direct(*d, grid)
            if k:
                order = np.argsort(-ref.astype(np.float32), kind="stable")[:k]
                np.testing.assert_array_equal(r.top_index, order)
                np.testing.assert_allclose(r.top_power, ref[order], atol=1e-6)
            else:
                np.testing.assert_allclose(r.power, ref, atol=1e-6)


def test_graph_entries_are_bounded_per_slot(serve):
    rng = np.random.default_rng(12)
    srv = serve.GLSServer(hbm_budget=1 << 24, slots=1, max_batch_curves=2,
                          max_batch_points=300, graphs_per_slot=2)
    used_before = srv.arena.used
    # each shape twice in a row (eager, then captured); the third capture
    # evicts the least recently used; a shape seen before is captured again
    # at once after its eviction
    for nf in (200, 200, 210, 210, 220, 220, 200):
        grid = 0.05 + 0.01 * np.arange(nf)
        for _ in range(2):
            m = int(rng.integers(20, 60))
            t = np.sort(rng.uniform(0, 20, m))
            srv.submit(t, np.sin(t) + rng.normal(0, 0.2, m), np.full(m, 0.2), grid)
        srv.drain()
    assert len(srv.slots[0].graphs) == 2 and srv.graph_captures == 4
    assert srv.graph_replays == 4 and srv.batches_launched == 7
    assert srv.arena.used > used_before
    assert all(r.power is not None for r in srv.completed)


def test_batch_cap_recovers_after_a_batch_that_did_not_fit(serve):
    """A batch that does not fit an idle server halves the batch size for the
    next attempt; every batch that launches doubles it back, up to
    max_batch_evals, so one tight moment does not shrink batches for good."""
    data, freqs = _requests(24)
    srv = serve.GLSServer(hbm_budget=1 << 22, slots=1, max_batch_evals=4000)
    real = srv._launch
    sizes = []

    def tight(slot, reqs):
        ev = sum(r.nf for r in reqs)
        if ev > 1500 and len(reqs) > 1:
            return False          # as if the arena could not hold it
        sizes.append(ev)
        return real(slot, reqs)

    srv._launch = tight
    first = [srv.submit(*d, f) for d, f in zip(data, freqs)]
    srv.drain()
    assert max(sizes) <= 1500 and srv.max_batch_evals == 4000
    for r, d, f in zip(first, data, freqs):
        np.testing.assert_allclose(r.power, og.gls_direct(*d, f), atol=1e-6)
    srv._launch = real
    sizes.clear()
    second = [srv.submit(*d, f) for d, f in zip(data, freqs)]
    n0 = srv.batches_launched
    srv.drain()
    assert srv._evals_cap == 4000
    assert srv.batches_launched - n0 <= 4     # batches of up to 4000 evaluations again
    for r, d, f in zip(second, data, freqs):
        np.testing.assert_allclose(r.power, og.gls_direct(*d, f), atol=1e-6)


def test_peaks_resolve_after_release(serve):
    """Releasing a zero-copy top-k result keeps copies of its k peaks; their
    frequencies and false-alarm probabilities are computed on first access,
    after the release, from those copies (not at release time, which the
    serving loop would pay for every request)."""
    from cuvarbase_amd.lombscargle import fap_baluev

    data, freqs = _requests(6, seed=21)
    srv = serve.GLSServer(hbm_budget=1 << 22, slots=2, max_batch_evals=700, zero_copy=True,
                          result_buffer_bytes=1 << 14)
    reqs = [srv.submit(*d, f, k=3) for d, f in zip(data, freqs)]
    srv.drain()
    for r in reqs:
        assert r._fap is None and r._top_freq is None
        r.release()
    assert srv.host_arena.used == 0
    for r, (t, y, dy), f in zip(reqs, data, freqs):
        P = og.gls_direct(t, y, dy, f)
        order = np.argsort(-P.astype(np.float32), kind="stable")[:3]
        np.testing.assert_array_equal(r.top_index, order)
        np.testing.assert_allclose(r.top_freq, f[order], rtol=1e-12)
        want = fap_baluev(t - np.mean(t), dy, np.asarray(r.top_power, np.float64), f[-1])
        np.testing.assert_allclose(r.fap, want, rtol=1e-12)
