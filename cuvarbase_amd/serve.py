"""Continuous-batching GLS server on one GPU.

Requests (one light curve + a linear frequency grid each) are queued as they
arrive.  Whenever one of ``slots`` batch slots is free, the server coalesces
the queued requests (FIFO, up to ``max_batch_evals`` curve x frequency
evaluations) into one planned batch, sub-allocates its device memory from a
resident arena (admission by HBM budget: a batch that does not fit waits for
memory to be released), and launches it on the slot's stream: H2D of the raw
(t, y, dy) -> prep -> passes -> D2H into the slot's pinned buffer.  Batches of
different slots overlap on the device; requests that arrive meanwhile join the
next batch.  ``poll`` retires finished batches (event query, no blocking) and
launches new ones.

The per-curve kernel parameters (grid size, Gaussian width) depend only on the
curve's own nf and on (sigma, m), never on the batch, so a request's result is
the same alone and in a mix up to the order of floating-point additions in the
LDS spreading.
"""
from __future__ import annotations

import bisect
import ctypes as C
import itertools
import time
from collections import deque

import numpy as np

from . import _lib
from ._device import DeviceArray, Event, Stream, pinned_empty, _handle
from .engine import Plan, linear_grid
from .tuning import choose_nfft_params


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


class Request:
    """One light curve submitted to the server."""

    _ids = itertools.count()

    def __init__(self, t, y, dy, freqs, arrival=None):
        self.id = next(self._ids)
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.dy = np.ascontiguousarray(dy, dtype=np.float64)
        self.f0, self.df, self.nf = linear_grid(freqs)
        self.arrival = time.perf_counter() if arrival is None else float(arrival)
        self.done_at = None
        self.power = None

    @property
    def latency(self):
        return None if self.done_at is None else self.done_at - self.arrival

    @property
    def done(self):
        return self.power is not None


class _Slot:
    def __init__(self):
        self.stream = Stream()
        self.event = Event()
        self.h_in = None      # pinned (3, max_obs) f64
        self.h_out = None     # pinned f32
        self.batch = None

    def staging(self, nobs, nout):
        if self.h_in is None or self.h_in.shape[1] < nobs:
            self.h_in = pinned_empty((3, max(nobs, 1024)), np.float64)
        if self.h_out is None or self.h_out.size < nout:
            self.h_out = pinned_empty(max(nout, 1024), np.float32)
        return self.h_in, self.h_out


class _Batch:
    def __init__(self, reqs, plan, blocks, nobs, nout, offsets):
        self.reqs = reqs
        self.plan = plan
        self.blocks = blocks
        self.nobs = nobs
        self.nout = nout
        self.offsets = offsets
        self.t_launch = time.perf_counter()


class GLSServer:
    """Continuous-batching server.

    Parameters
    ----------
    hbm_budget : bytes of device memory the server may use (resident arena)
    slots : number of batches that may be in flight at once (one stream each)
    max_batch_evals : cap on curve x frequency evaluations per batch
    sigma, m : NFFT parameters (default: tuned for |dP| <= ``tol``)
    """

    def __init__(self, hbm_budget=16e9, slots=3, max_batch_evals=3e8, max_batch_curves=8192,
                 sigma=None, m=None, tol=1e-4, mode=_lib.MODE_FLOATING_MEAN):
        _lib.init(0)
        s0, m0 = choose_nfft_params(tol)
        self.sigma = float(sigma if sigma is not None else s0)
        self.m = int(m if m is not None else m0)
        self.mode = mode
        self.max_batch_evals = int(max_batch_evals)
        self.max_batch_curves = int(max_batch_curves)
        self.arena = DeviceArena(hbm_budget)
        self.slots = [_Slot() for _ in range(slots)]
        self.queue = deque()
        self.completed = []
        self.batches_launched = 0
        self.evals_done = 0

    # -- client side ---------------------------------------------------------
    def submit(self, t, y, dy, freqs, arrival=None):
        r = Request(t, y, dy, freqs, arrival=arrival)
        self.queue.append(r)
        return r

    # -- server loop -----------------------------------------------------------
    def _retire(self, slot):
        b = slot.batch
        now = time.perf_counter()
        out = slot.h_out
        for r, o in zip(b.reqs, b.offsets):
            r.power = np.array(out[o:o + r.nf])
            r.done_at = now
            self.completed.append(r)
            self.evals_done += r.nf
        for blk in b.blocks:
            self.arena.release(blk)
        b.plan.destroy()
        slot.batch = None

    def _take_batch(self):
        reqs, ev = [], 0
        while self.queue and len(reqs) < self.max_batch_curves:
            r = self.queue[0]
            if reqs and ev + r.nf > self.max_batch_evals:
                break
            reqs.append(self.queue.popleft())
            ev += r.nf
        return reqs

    def _launch(self, slot, reqs):
        curves, obs, out, offsets = [], 0, 0, []
        for r in reqs:
            curves.append(_lib.gls_curve(obs_offset=obs, n_obs=len(r.t), nf=r.nf, f0=r.f0,
                                         df=r.df, out_offset=out, ybar=0.0, yy=1.0))
            offsets.append(out)
            obs += len(r.t)
            out += r.nf
        plan = Plan.lomb(curves, self.sigma, self.m, self.mode, 0.0, 0.0, _lib.INPUT_RAW,
                         _lib.METHOD_NFFT)
        sizes = [24 * max(obs, 1), 4 * max(out, 1), plan.workspace_bytes]
        blocks = []
        for sz in sizes:
            blk = self.arena.alloc(sz)
            if blk is None:
                for b in blocks:
                    self.arena.release(b)
                plan.destroy()
                return False
            blocks.append(blk)
        h_in, h_out = slot.staging(obs, out)
        o = 0
        for r in reqs:
            n = len(r.t)
            h_in[0, o:o + n] = r.t
            h_in[1, o:o + n] = r.y
            h_in[2, o:o + n] = r.dy
            o += n
        s = slot.stream.handle
        p_in = self.arena.ptr(blocks[0])
        row = 8 * max(obs, 1)
        for k in range(3):
            _lib.call("gls_memcpy_h2d_async", p_in + k * row, h_in[k].ctypes.data,
                      C.c_size_t(8 * obs), s)
        _lib.call("gls_lomb_execute", plan.handle, p_in, p_in + row, p_in + 2 * row,
                  self.arena.ptr(blocks[1]), self.arena.ptr(blocks[2]), s)
        _lib.call("gls_memcpy_d2h_async", h_out.ctypes.data, self.arena.ptr(blocks[1]),
                  C.c_size_t(4 * out), s)
        slot.event.record(slot.stream)
        slot.batch = _Batch(reqs, plan, blocks, obs, out, offsets)
        self.batches_launched += 1
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
            if not self._launch(slot, reqs):
                # not enough device memory now: requeue at the front and wait
                # for in-flight batches to release theirs
                self.queue.extendleft(reversed(reqs))
                if not any(s.batch is not None for s in self.slots):
                    if len(reqs) == 1:
                        self.queue.popleft()
                        raise MemoryError(
                            f"request {reqs[0].id} needs more device memory than the "
                            f"server budget ({self.arena.nbytes} bytes)")
                    # retry with a smaller batch next time
                    self.max_batch_evals = max(1, self.max_batch_evals // 2)
                break
        return len(self.completed) - before

    @property
    def busy(self):
        return bool(self.queue) or any(s.batch is not None for s in self.slots)

    def drain(self, sleep=5e-5):
        while self.busy:
            if self.poll() == 0:
                time.sleep(sleep)

    # -- measurement -------------------------------------------------------------
    def serve_stream(self, arrivals, data, freqs, sleep=5e-5):
        """Replay a request stream in real time: request i arrives arrivals[i]
        seconds after the start.  Returns a stats dict."""
        t0 = time.perf_counter()
        reqs = []
        i, n = 0, len(data)
        while i < n or self.busy:
            now = time.perf_counter() - t0
            while i < n and arrivals[i] <= now:
                t, y, dy = data[i]
                reqs.append(self.submit(t, y, dy, freqs[i], arrival=t0 + arrivals[i]))
                i += 1
            if self.poll() == 0 and (i >= n or arrivals[i] > time.perf_counter() - t0):
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
            "batches": self.batches_launched,
            "requests_obj": reqs,
        }
