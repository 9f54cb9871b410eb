"""Continuous-batching GLS server on one GPU.

Requests (one light curve + a linear frequency grid, or an explicit list of
frequencies, each) are queued as they arrive.  Whenever one of ``slots`` batch
slots is free, the server coalesces the queued requests of the oldest
request's class (same method, GLS mode and result kind; FIFO, up to
``max_batch_evals`` curve x frequency evaluations) into one planned batch, sub-allocates its device memory from a
resident arena (admission by HBM budget: a batch that does not fit waits for
memory to be released), and launches it on the slot's stream: H2D of the raw
(t, y, dy) -> prep -> passes [-> top-k] -> D2H into the slot's pinned buffer
(the whole periodogram, or only k peaks per curve for top-k requests).  Batches of
different slots overlap on the device; requests that arrive meanwhile join the
next batch.  ``poll`` retires finished batches (event query, no blocking) and
launches new ones.

The per-curve kernel parameters (grid size, Gaussian width) depend only on the
curve's own nf and on (sigma, m), never on the batch, and the spreading adds in
fixed point, so a request's result is bitwise the same alone and in a mix.
"""
from __future__ import annotations

import bisect
import contextlib
import ctypes as C
import gc
import itertools
import time
from collections import OrderedDict, deque

import numpy as np

from . import _lib
from ._device import DeviceArray, Event, Stream, pinned_empty, _handle
from .engine import Graph, Plan, linear_grid
from .tuning import choose_nfft_config
from . import wire


class FreeList:
    """First-fit allocator bookkeeping over [0, nbytes) with coalescing."""

    ALIGN = 256

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self._free = [(0, self.nbytes)]       # sorted (offset, size)
        self.used = 0

    def alloc(self, nbytes):
        need = max((int(nbytes) + self.ALIGN - 1) // self.ALIGN * self.ALIGN, self.ALIGN)
        for i, (off, size) in enumerate(self._free):
            if size >= need:
                if size == need:
                    self._free.pop(i)
                else:
                    self._free[i] = (off + need, size - need)
                self.used += need
                return off, need
        return None

    def release(self, block):
        off, size = block
        self.used -= size
        i = bisect.bisect_left(self._free, (off, 0))
        self._free.insert(i, (off, size))
        if i + 1 < len(self._free) and off + size == self._free[i + 1][0]:
            self._free[i] = (off, size + self._free[i + 1][1])
            self._free.pop(i + 1)
        if i > 0 and self._free[i - 1][0] + self._free[i - 1][1] == self._free[i][0]:
            self._free[i - 1] = (self._free[i - 1][0], self._free[i - 1][1] + self._free[i][1])
            self._free.pop(i)

    @property
    def largest_free(self):
        return max((s for _, s in self._free), default=0)


class DeviceArena(FreeList):
    """FreeList over one resident device buffer: serving never calls
    hipMalloc/hipFree after start-up."""

    def __init__(self, nbytes):
        super().__init__(nbytes)
        self.buf = DeviceArray(self.nbytes, np.uint8)

    def ptr(self, block):
        return self.buf.ptr + block[0]


class HostArena(FreeList):
    """FreeList over one page-locked host buffer: D2H copies land in it
    directly and results are handed out as views (no per-batch host
    allocation, no copy into fresh pages)."""

    def __init__(self, nbytes):
        super().__init__(nbytes)
        self.buf = pinned_empty(self.nbytes, np.uint8)

    def view(self, block, dtype, count, offset=0):
        start = block[0] + offset
        return self.buf[start:start + count * np.dtype(dtype).itemsize].view(dtype)

    def ptr(self, block):
        return self.buf.ctypes.data + block[0]


class _HostHold:
    """A result block shared by the requests of one batch."""

    def __init__(self, arena, block, count):
        self.arena, self.block, self.count = arena, block, count

    def drop(self):
        self.count -= 1
        if self.count == 0:
            self.arena.release(self.block)


class Request:
    """One light curve submitted to the server.

    ``freqs`` is a linear grid (evaluated by the NFFT path) unless
    ``custom_freqs`` is set (direct sums at exactly those frequencies).
    ``k > 0`` asks for the k highest peaks (power, index, frequency and
    Baluev false-alarm probability) instead of the whole periodogram."""

    _ids = itertools.count()

    def __init__(self, t, y, dy, freqs, arrival=None, custom_freqs=False, k=0,
                 mode=None, rid=None):
        self.id = next(self._ids) if rid is None else int(rid)
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.dy = np.ascontiguousarray(dy, dtype=np.float64)
        if custom_freqs:
            self.freqs = np.ascontiguousarray(freqs, dtype=np.float64)
            if self.freqs.ndim != 1 or len(self.freqs) < 1:
                raise ValueError("custom frequencies must be a non-empty 1-D array")
            self.f0, self.df, self.nf = 0.0, 1.0, len(self.freqs)
        else:
            self.freqs = None
            self.f0, self.df, self.nf = linear_grid(freqs)
        self.k = int(k)
        if self.k < 0 or self.k > 64:
            raise ValueError("k must be in [0, 64]")
        self.mode = mode
        self.arrival = time.perf_counter() if arrival is None else float(arrival)
        self.done_at = None
        self.power = None
        self.top_power = self.top_index = None
        self._top_freq = self._fap = None
        self._hold = None

    @property
    def top_freq(self):
        """Frequencies of the top-k peaks (NaN where fewer than k exist)."""
        if self._top_freq is None and self.top_index is not None:
            self._top_freq = self.frequency(self.top_index)
        return self._top_freq

    @property
    def fap(self):
        """Baluev false-alarm probability of each top-k peak, computed on first
        access (outside the serving loop)."""
        if self._fap is None and self.top_power is not None:
            self._fap = _fap(self, self.top_power, self.top_index)
        return self._fap

    def release(self):
        """Give a zero-copy result's host memory back to the server; the
        result arrays must not be used afterwards."""
        h, self._hold = self._hold, None
        if h is not None:
            if self.top_power is not None:
                # the peaks outlive their memory as copies (k values each);
                # their frequencies and false-alarm probabilities are still
                # computed from these on first access
                self.top_power = np.array(self.top_power)
                self.top_index = np.array(self.top_index)
            self.power = None
            h.drop()

    @property
    def latency(self):
        return None if self.done_at is None else self.done_at - self.arrival

    @property
    def done(self):
        return self.done_at is not None

    def frequency(self, idx):
        idx = np.asarray(idx)
        if self.freqs is not None:
            return np.where(idx >= 0, self.freqs[np.maximum(idx, 0)], np.nan)
        return np.where(idx >= 0, self.f0 + self.df * idx, np.nan)


class _Slot:
    def __init__(self):
        self.stream = Stream()
        self.event = Event()
        self.h_in = None      # pinned f64: t, y, dy of a batch, one after another
        self.h_out = None     # pinned f32
        self.h_freq = None    # pinned f64 (custom frequency lists)
        self.h_idx = None     # pinned i32 (top-k indices)
        self.batch = None
        self.graphs = OrderedDict()   # repeated batch shape -> _GraphEntry

    def staging(self, nobs, nout, nfreq=0, nidx=0):
        # page-locked staging grows geometrically and is kept for the slot's life
        def grow(cur, need):
            return max(need, 1024, 0 if cur is None else 2 * len(cur))

        if self.h_in is None or self.h_in.size < 3 * nobs:
            self.h_in = pinned_empty(grow(self.h_in, 3 * nobs), np.float64)
        if nout and (self.h_out is None or self.h_out.size < nout):
            self.h_out = pinned_empty(grow(self.h_out, nout), np.float32)
        if nfreq and (self.h_freq is None or self.h_freq.size < nfreq):
            self.h_freq = pinned_empty(grow(self.h_freq, nfreq), np.float64)
        if nidx and (self.h_idx is None or self.h_idx.size < nidx):
            self.h_idx = pinned_empty(grow(self.h_idx, nidx), np.int32)
        return self.h_in, self.h_out


class _Batch:
    def __init__(self, reqs, plan, blocks, nobs, nout, offsets, k=0, hblock=None):
        self.reqs = reqs
        self.plan = plan
        self.blocks = blocks
        self.nobs = nobs
        self.nout = nout
        self.offsets = offsets
        self.k = k
        self.hblock = hblock      # result block in the host arena (zero-copy mode)
        self.t_launch = time.perf_counter()


class _GraphEntry:
    """A captured execute of one repeated batch shape on one slot, with the
    device blocks it reads and writes at fixed addresses: the inputs (rows of
    t, y, dy, each ``cap`` points long), the powers, the workspace and the
    top-k output.  Later batches of the shape rebind the plan to their curves
    and replay the graph."""

    def __init__(self, plan, graph, blocks, cap):
        self.plan, self.graph, self.blocks, self.cap = plan, graph, blocks, cap
        self.uses = 0


class GLSServer:
    """Continuous-batching server.

    Parameters
    ----------
    hbm_budget : bytes of device memory the server may use (resident arena)
    slots : number of batches that may be in flight at once (one stream each)
    max_batch_evals : cap on curve x frequency evaluations per batch
    max_batch_points : cap on observations per batch; each slot's page-locked
        input staging (24 B per point) is allocated for it at start-up (a
        single larger request still runs, growing its slot's staging)
    sigma, m, window, window_beta : NFFT parameters (default: tuned for
        |dP| <= ``tol``)
    spread : ``_lib.SPREAD_FIXED`` (the default) accumulates the spreading in
        fixed point, so a request's result is bitwise the same alone and in any
        mix; ``SPREAD_FLOAT`` uses FP32 atomics (equal to ~1e-7, the order of
        FP32 additions)
    graph_batches : batches whose requests share one linear frequency grid
        (and have at most 16,384 points each) run as replays of a captured
        execute kept per (shape, slot) -- the plan is rebound to the new
        curves and only its descriptors and the inputs are uploaded.  A shape
        is captured the second time a batch of it runs (the first runs
        eagerly: a capture costs more host time than an eager launch, and
        shapes seen once -- a lone request on its own grid -- would never
        repay it); at most ``graphs_per_slot`` shapes are kept per slot
        (least recently used first out), their device blocks held in the
        arena meanwhile
    zero_copy : results are views into a page-locked result arena of
        ``result_buffer_bytes`` (default 4 B x max_batch_evals x (slots + 1))
        that the D2H copies write directly; each request's memory returns to
        the arena on ``Request.release()``.  Otherwise results are copied into
        arrays the requests own.
    """

    def __init__(self, hbm_budget=16e9, slots=3, max_batch_evals=3e8, max_batch_curves=8192,
                 sigma=None, m=None, tol=1e-4, mode=_lib.MODE_FLOATING_MEAN, zero_copy=False,
                 result_buffer_bytes=None, window=None, window_beta=0.0,
                 max_batch_points=1 << 20, graph_batches=True, graphs_per_slot=2,
                 spread=_lib.SPREAD_FIXED):
        _lib.init(0)
        cfg = choose_nfft_config(tol)
        if sigma is not None or m is not None or window is not None:
            cfg = dict(sigma=sigma if sigma is not None else cfg["sigma"],
                       m=m if m is not None else cfg["m"],
                       window=window if window is not None else cfg["window"],
                       window_beta=window_beta)
        self.sigma = float(cfg["sigma"])
        self.m = int(cfg["m"])
        self.window = int(cfg["window"])
        self.window_beta = float(cfg["window_beta"])
        self.mode = mode
        self.max_batch_evals = int(max_batch_evals)
        # batch size in evaluations now: halved while a batch does not fit
        # the free memory of an idle server, doubled back (up to
        # max_batch_evals) by every batch that launches
        self._evals_cap = self.max_batch_evals
        self.max_batch_curves = int(max_batch_curves)
        self.max_batch_points = max(int(max_batch_points), 1)
        self.graph_batches = bool(graph_batches)
        self.spread = int(spread)
        self.graphs_per_slot = max(int(graphs_per_slot), 1)
        self.graph_captures = 0
        self.graph_replays = 0
        self.launch_host_s = []             # host time of each batch launch (seconds)
        self.slow_launches = []             # step breakdown of launches over 2 ms
        self._last_launch = None
        self._shapes_seen = OrderedDict()   # batch shapes run at least once (bounded)
        self.arena = DeviceArena(hbm_budget)
        self.zero_copy = bool(zero_copy)
        self.host_arena = None
        if self.zero_copy:
            nb = result_buffer_bytes or 4 * self.max_batch_evals * (slots + 1)
            self.host_arena = HostArena(int(nb))
        self.slots = [_Slot() for _ in range(slots)]
        for slot in self.slots:
            slot.staging(self.max_batch_points, 0)
        self.queue = deque()
        self.completed = []
        self.batches_launched = 0
        self.evals_done = 0
        self._warm_up()

    def _warm_up(self):
        """One tiny request through the whole path at start-up: the first plan
        of a configuration (its deconvolution table) and the process's first
        kernel launches cost tens of milliseconds, which would otherwise land
        in the first real request's latency."""
        t = np.linspace(0.0, 10.0, 64)
        r = self.submit(t, np.sin(t), np.ones_like(t), 0.1 + 0.05 * np.arange(64))
        try:
            self.drain()
        except MemoryError:
            # a budget too small for even this request: start cold
            self.queue.clear()
        r.release()
        for slot in self.slots:
            for key in list(slot.graphs):
                self._drop_graph(slot, key)
        self.completed.clear()
        self.batches_launched = 0
        self.evals_done = 0
        self.graph_captures = self.graph_replays = 0
        self.launch_host_s = []
        self.slow_launches = []
        self._shapes_seen.clear()

    # -- client side ---------------------------------------------------------
    def submit(self, t, y, dy, freqs, arrival=None, custom_freqs=False, k=0, mode=None,
               rid=None):
        r = Request(t, y, dy, freqs, arrival=arrival, custom_freqs=custom_freqs, k=k,
                    mode=mode, rid=rid)
        self.queue.append(r)
        return r

    # -- server loop -----------------------------------------------------------
    def _retire(self, slot):
        b = slot.batch
        now = time.perf_counter()
        nreq = len(b.reqs)
        if b.hblock is not None:
            hold = _HostHold(self.host_arena, b.hblock, nreq)
            if b.k:
                out = self.host_arena.view(b.hblock, np.float32, nreq * b.k)
                idx = self.host_arena.view(b.hblock, np.int32, nreq * b.k, 4 * nreq * b.k)
            else:
                out = self.host_arena.view(b.hblock, np.float32, b.nout)
        else:
            hold = None
            out = slot.h_out
            idx = slot.h_idx
        for i, (r, o) in enumerate(zip(b.reqs, b.offsets)):
            if b.k:
                k = r.k
                r.top_power = out[i * b.k:i * b.k + k]
                r.top_index = idx[i * b.k:i * b.k + k]
                if hold is None:
                    r.top_power = np.array(r.top_power)
                    r.top_index = np.array(r.top_index)
            else:
                r.power = out[o:o + r.nf] if hold is not None else np.array(out[o:o + r.nf])
            r._hold = hold
            r.done_at = now
            self.completed.append(r)
            self.evals_done += r.nf
        for blk in b.blocks:
            self.arena.release(blk)
        if b.plan is not None:
            b.plan.destroy()
        slot.batch = None

    def _class(self, r):
        return (r.freqs is not None, self.mode if r.mode is None else r.mode, r.k > 0)

    def _take_batch(self):
        """FIFO head's class; later requests of other classes keep their place."""
        if not self.queue:
            return []
        cls = self._class(self.queue[0])
        reqs, rest, ev, pts = [], [], 0, 0
        while self.queue and len(reqs) < self.max_batch_curves:
            r = self.queue[0]
            if self._class(r) != cls:
                rest.append(self.queue.popleft())
                continue
            if reqs and (ev + r.nf > self._evals_cap
                         or pts + len(r.t) > self.max_batch_points):
                break
            reqs.append(self.queue.popleft())
            ev += r.nf
            pts += len(r.t)
        self.queue.extendleft(reversed(rest))
        return reqs

    def _graph_key(self, reqs, mode, k):
        """Key of a repeatable batch shape, or None: every request on one
        linear grid and within the one-block preparation limit."""
        if not self.graph_batches:
            return None
        r0 = reqs[0]
        if r0.freqs is not None:
            return None
        grid = (r0.f0, r0.df, r0.nf)
        for r in reqs:
            if r.freqs is not None or (r.f0, r.df, r.nf) != grid or len(r.t) > 16384:
                return None
        return (mode, k, grid, len(reqs))

    def _drop_graph(self, slot, key):
        entry = slot.graphs.pop(key)
        entry.graph.destroy()
        entry.plan.destroy()
        for blk in entry.blocks:
            self.arena.release(blk)

    SHAPES_REMEMBERED = 1024

    def _graph_entry(self, slot, key, curves, obs):
        """The slot's captured execute for ``key``, bound to ``curves``
        (captured when the shape runs for the second time), or None for the
        first batch of a shape, if the arena cannot hold a new capture or if
        the batch does not fit the captured one."""
        entry = slot.graphs.get(key)
        if entry is not None:
            try:
                entry.plan.rebind(curves)
            except _lib.GLSError:
                return None
            slot.graphs.move_to_end(key)
            return entry
        if key not in self._shapes_seen:
            self._shapes_seen[key] = True
            if len(self._shapes_seen) > self.SHAPES_REMEMBERED:
                self._shapes_seen.popitem(last=False)
            return None
        self._shapes_seen.move_to_end(key)
        mode, k, (f0, df, nf), nreq = key
        cap = max(self.max_batch_points, obs)
        plan = Plan.lomb(curves, self.sigma, self.m, mode, 0.0, 0.0, _lib.INPUT_RAW,
                         _lib.METHOD_NFFT, window=self.window, window_beta=self.window_beta,
                         point_capacity=cap, spread=self.spread)
        while len(slot.graphs) >= self.graphs_per_slot:
            self._drop_graph(slot, next(iter(slot.graphs)))
        sizes = [24 * cap, 4 * nreq * nf, plan.workspace_bytes] + ([8 * nreq * k] if k else [])
        blocks = []
        for sz in sizes:
            blk = self.arena.alloc(sz)
            if blk is None:
                for b in blocks:
                    self.arena.release(b)
                plan.destroy()
                return None
            blocks.append(blk)
        p_in = self.arena.ptr(blocks[0])
        row = 8 * cap
        g = C.c_void_p()
        _lib.call("gls_plan_graph_create", plan.handle, p_in, p_in + row, p_in + 2 * row,
                  self.arena.ptr(blocks[1]), C.c_int32(0), self.arena.ptr(blocks[2]),
                  slot.stream.handle, C.byref(g))
        entry = _GraphEntry(plan, Graph(g.value), blocks, cap)
        slot.graphs[key] = entry
        self.graph_captures += 1
        return entry

    def _launch(self, slot, reqs):
        custom, mode, topk = self._class(reqs[0])
        curves, obs, out, offsets = [], 0, 0, []
        for r in reqs:
            curves.append(_lib.gls_curve(obs_offset=obs, n_obs=len(r.t), nf=r.nf, f0=r.f0,
                                         df=r.df, out_offset=out, ybar=0.0, yy=1.0))
            offsets.append(out)
            obs += len(r.t)
            out += r.nf
        k = max(r.k for r in reqs) if topk else 0
        nreq = len(reqs)
        marks = [time.perf_counter()]
        key = None if custom else self._graph_key(reqs, mode, k)
        entry = self._graph_entry(slot, key, curves, obs) if key is not None else None
        marks.append(time.perf_counter())
        if entry is not None:
            # a replay: the entry's blocks stay with it after the batch
            plan, blocks, dev = entry.plan, [], entry.blocks
            row = 8 * entry.cap
        else:
            plan = Plan.lomb(curves, self.sigma, self.m, mode, 0.0, 0.0, _lib.INPUT_RAW,
                             _lib.METHOD_DIRECT_CUSTOM if custom else _lib.METHOD_NFFT,
                             window=self.window, window_beta=self.window_beta,
                             spread=_lib.SPREAD_FIXED if custom else self.spread)
            sizes = [24 * max(obs, 1), 4 * max(out, 1), plan.workspace_bytes]
            sizes.append(8 * nreq * k if k else 0)
            if custom:
                sizes.append(8 * max(out, 1))
            blocks = []
            for sz in sizes:
                blk = self.arena.alloc(sz)
                if blk is None:
                    for b in blocks:
                        self.arena.release(b)
                    plan.destroy()
                    return False
                blocks.append(blk)
            dev = blocks
            row = 8 * max(obs, 1)
        marks.append(time.perf_counter())
        hblock = None
        if self.zero_copy:
            hblock = self.host_arena.alloc(8 * nreq * k if k else 4 * max(out, 1))
            if hblock is None:
                for b in blocks:
                    self.arena.release(b)
                if entry is None:
                    plan.destroy()
                return False
        h_in, h_out = slot.staging(obs, 0 if self.zero_copy else (nreq * k if k else out),
                                   out if custom else 0, 0 if self.zero_copy else nreq * k)
        # rows t, y, dy of the batch back to back
        rows = h_in[:3 * obs].reshape(3, obs)
        o = 0
        for r in reqs:
            n = len(r.t)
            rows[0, o:o + n] = r.t
            rows[1, o:o + n] = r.y
            rows[2, o:o + n] = r.dy
            o += n
        marks.append(time.perf_counter())
        s = slot.stream.handle
        p_in, p_pow, p_ws = (self.arena.ptr(b) for b in dev[:3])
        if entry is not None:
            # rows land at the captured row stride: three copies; then the
            # rebound descriptors, then the replay
            for q in range(3):
                _copy_in(p_in + q * row, h_in.ctypes.data + q * 8 * obs, 8 * obs, s)
            if entry.uses:
                entry.graph.update(plan, slot.stream)
            entry.graph.launch(slot.stream)
            entry.uses += 1
            self.graph_replays += 1
        else:
            _copy_in(p_in, h_in.ctypes.data, 24 * obs, s)
            marks.append(time.perf_counter())
            if custom:
                for r, o in zip(reqs, offsets):
                    slot.h_freq[o:o + r.nf] = r.freqs
                p_fr = self.arena.ptr(dev[4])
                _copy_in(p_fr, slot.h_freq.ctypes.data, 8 * out, s)
                _lib.call("gls_lomb_execute_custom", plan.handle, p_in, p_in + row,
                          p_in + 2 * row, p_fr, p_pow, C.c_int32(0), p_ws, s)
            else:
                _lib.call("gls_lomb_execute", plan.handle, p_in, p_in + row, p_in + 2 * row,
                          p_pow, p_ws, s)
            marks.append(time.perf_counter())
        if k:
            p_top = self.arena.ptr(dev[3])
            _lib.call("gls_topk", plan.handle, p_ws, p_pow, None, C.c_int32(k), p_top,
                      p_top + 4 * nreq * k, s)
            if hblock is not None:
                _copy_out(self.host_arena.ptr(hblock), p_top, 8 * nreq * k, s)
            else:
                _copy_out(h_out.ctypes.data, p_top, 4 * nreq * k, s)
                _copy_out(slot.h_idx.ctypes.data, p_top + 4 * nreq * k, 4 * nreq * k, s)
        else:
            dst = self.host_arena.ptr(hblock) if hblock is not None else h_out.ctypes.data
            _copy_out(dst, p_pow, 4 * out, s)
        marks.append(time.perf_counter())
        slot.event.record(slot.stream)
        marks.append(time.perf_counter())
        slot.batch = _Batch(reqs, None if entry is not None else plan, blocks, obs, out,
                            offsets, k=k, hblock=hblock)
        self.batches_launched += 1
        # host time of the launch's steps (ms): graph lookup or capture, plan
        # and memory, staging, copies and kernel launches
        d = [float(x) for x in np.diff(marks) * 1e3]
        info = {"requests": nreq, "points": obs, "evals": out,
                "graph": "replay" if entry is not None else "eager",
                "graph_ms": d[0], "plan_and_memory_ms": d[1], "staging_ms": d[2]}
        if entry is None:
            # eager: input copy, execute, result copies (and top-k), event record
            info.update(h2d_ms=d[3], execute_ms=d[4], results_ms=d[5], event_ms=d[6])
        else:
            info.update(replay_ms=d[3], event_ms=d[4])
        self._last_launch = info
        return True

    def poll(self):
        """Retire finished batches and launch queued work on free slots.
        Returns the number of requests completed by this call."""
        before = len(self.completed)
        for slot in self.slots:
            if slot.batch is not None and slot.event.query():
                self._retire(slot)
        for slot in self.slots:
            if not self.queue:
                break
            if slot.batch is not None:
                continue
            reqs = self._take_batch()
            t_l = time.perf_counter()
            ok = self._launch(slot, reqs)
            t_l = time.perf_counter() - t_l
            self.launch_host_s.append(t_l)
            if ok and t_l > 2e-3 and len(self.slow_launches) < 50:
                self.slow_launches.append(dict(self._last_launch, total_ms=t_l * 1e3,
                                               at_launch=len(self.launch_host_s) - 1))
            if not ok:
                # not enough device memory now: requeue at the front and wait
                # for in-flight batches to release theirs
                self.queue.extendleft(reversed(reqs))
                if not any(s.batch is not None for s in self.slots):
                    if len(reqs) == 1:
                        self.queue.popleft()
                        if self.zero_copy and self.host_arena.used > 0:
                            raise MemoryError(
                                f"request {reqs[0].id}: the result arena is held by "
                                "unreleased results (call Request.release())")
                        raise MemoryError(
                            f"request {reqs[0].id} needs more device or result memory than "
                            f"the server budget ({self.arena.nbytes} bytes)")
                    # retry with a smaller batch next time
                    self._evals_cap = max(1, self._evals_cap // 2)
                break
            self._evals_cap = min(self.max_batch_evals, 2 * self._evals_cap)
        return len(self.completed) - before

    @property
    def busy(self):
        return bool(self.queue) or any(s.batch is not None for s in self.slots)

    def drain(self, sleep=5e-5):
        while self.busy:
            if self.poll() == 0:
                time.sleep(sleep)

    @staticmethod
    @contextlib.contextmanager
    def _gc_frozen():
        # objects that exist when a serving loop starts (the caller's data,
        # the server) leave the collector's generations, so its occasional
        # full collections stay short inside the loop
        froze = gc.get_freeze_count() == 0
        if froze:
            gc.freeze()
        try:
            yield
        finally:
            if froze:
                gc.unfreeze()

    # -- measurement -------------------------------------------------------------
    def serve_stream(self, arrivals, data, freqs, sleep=5e-5, release=False, k=0):
        """Replay a request stream in real time: request i arrives arrivals[i]
        seconds after the start (asking for the k highest peaks when k > 0).
        Returns a stats dict.  ``release=True`` hands each result back
        (zero-copy mode) as soon as it completes, as a consumer that has
        taken it would."""
        reqs = []
        i, n = 0, len(data)
        nl0 = len(self.launch_host_s)
        ns0 = len(self.slow_launches)
        b0, c0, r0 = self.batches_launched, self.graph_captures, self.graph_replays
        with self._gc_frozen():
            t0 = time.perf_counter()
            while i < n or self.busy:
                now = time.perf_counter() - t0
                while i < n and arrivals[i] <= now:
                    t, y, dy = data[i]
                    reqs.append(self.submit(t, y, dy, freqs[i], arrival=t0 + arrivals[i], k=k))
                    i += 1
                done = self.poll()
                if release and done:
                    for r in self.completed[-done:]:
                        r.release()
                if done == 0 and (i >= n or arrivals[i] > time.perf_counter() - t0):
                    time.sleep(sleep)
            wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in reqs])
        evals = sum(r.nf for r in reqs)
        return {
            "requests": n,
            "evals": evals,
            "wall_s": wall,
            "evals_per_s": evals / wall,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "batches": self.batches_launched - b0,
            "graph_captures": self.graph_captures - c0,
            "graph_replays": self.graph_replays - r0,
            # host time of this stream's batch launches (plan, staging, copies,
            # kernel launches): the serving loop's own cost per batch
            "launch_host_ms": _quantiles_ms(self.launch_host_s[nl0:]),
            "slow_launches": self.slow_launches[ns0:],
            "requests_obj": reqs,
        }

    # -- record streams ----------------------------------------------------------
    def serve_records(self, reader, writer, sleep=5e-5):
        """Serve request records read from the binary stream ``reader`` (until
        EOF) and write one result record per request to ``writer``, in
        completion order.  Returns the number of results written."""
        written = 0

        def flush():
            nonlocal written
            for r in self.completed:
                writer.write(wire.encode_result(_result_record(r)))
                r.release()
                written += 1
            self.completed.clear()

        with self._gc_frozen():
            for rec in wire.iter_requests(reader):
                if rec.freqs is not None:
                    self.submit(rec.t, rec.y, rec.dy, rec.freqs, custom_freqs=True, k=rec.k,
                                mode=rec.mode, rid=rec.id)
                else:
                    if rec.nf < 1 or not (rec.df > 0):
                        raise wire.WireError(f"request {rec.id}: bad linear grid")
                    self.submit(rec.t, rec.y, rec.dy, rec.frequencies(), k=rec.k,
                                mode=rec.mode, rid=rec.id)
                self.poll()
                flush()
            while self.busy:
                if self.poll() == 0:
                    time.sleep(sleep)
                flush()
            flush()
        return written


# Copies between the page-locked staging / result memory and the arena: up to
# KERNEL_COPY_MAX bytes by a copy kernel (gls_memcpy_kernel_async), which is
# enqueued like any launch; the DMA path's hipMemcpyAsync was measured holding
# the serving thread for 7-9 ms now and then on such small copies.  Larger
# copies (bursts of large batches) keep the DMA engines' bandwidth.
KERNEL_COPY_MAX = 16 << 20


def _copy_in(dst, src, nbytes, stream):
    fn = "gls_memcpy_kernel_async" if nbytes <= KERNEL_COPY_MAX else "gls_memcpy_h2d_async"
    _lib.call(fn, dst, src, C.c_size_t(nbytes), stream)


def _copy_out(dst, src, nbytes, stream):
    fn = "gls_memcpy_kernel_async" if nbytes <= KERNEL_COPY_MAX else "gls_memcpy_d2h_async"
    _lib.call(fn, dst, src, C.c_size_t(nbytes), stream)


def _quantiles_ms(v):
    if not v:
        return None
    a = np.asarray(v) * 1e3
    return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99)),
            "max": float(a.max()), "over_1ms": int(np.sum(a > 1.0)), "count": int(a.size)}


def _fap(r, z, idx):
    """Baluev false-alarm probability of each peak (fap_baluev,
    cuvarbase/lombscargle.py:1115-1184, with fmax the grid's highest frequency)."""
    from .lombscargle import fap_baluev

    fmax = float(np.max(r.freqs)) if r.freqs is not None else r.f0 + r.df * (r.nf - 1)
    z = np.asarray(z, dtype=np.float64)
    out = np.full(len(z), np.nan)
    ok = np.asarray(idx) >= 0
    if ok.any():
        out[ok] = fap_baluev(r.t - np.mean(r.t), r.dy, z[ok], fmax)
    return out


def _result_record(r):
    lat = r.latency if r.latency is not None else float("nan")
    if r.top_power is not None:
        return wire.ResultRecord(r.id, top_power=r.top_power, top_index=r.top_index,
                                 top_freq=r.top_freq, fap=r.fap, latency=lat)
    return wire.ResultRecord(r.id, power=r.power, latency=lat)
