"""Batched GLS engine over the libgls C ABI.

A ``LombBatch`` packs many light curves into flat device arrays (observations
concatenated, one output segment per curve), plans them once
(``gls_lomb_plan_create``) and runs every curve of the batch in one launch
sequence.  ``NFFTBatch`` does the same for plain adjoint NFFTs.  The drop-in
classes in ``lombscargle`` / ``cunfft`` and the serving loop are built on these.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._device import DeviceArray, Stream, _handle, pinned_empty


def linear_grid(freqs):
    """(f0, df, nf) of a linear frequency grid (the reference reads
    df = freqs[1] - freqs[0], cuvarbase/lombscargle.py:623)."""
    freqs = np.asarray(freqs, dtype=np.float64)
    if freqs.ndim != 1 or len(freqs) < 2:
        raise ValueError("a frequency grid needs at least two frequencies")
    return float(freqs[0]), float(freqs[1] - freqs[0]), int(len(freqs))


def set_tiles(n1, p2, c1=0, cr=0):
    """Override the NFFT pass tiles of size bucket (n1, p2) for plans made
    from now on (0 keeps the library's choice for that pass; both 0 remove
    the override).  Raises GLSError if the library has no such kernel."""
    _lib.call("gls_tile_set", C.c_int32(n1), C.c_int32(p2), C.c_int32(c1), C.c_int32(cr))


def get_tiles(n1, p2):
    """(c1, cr) that a plan made now would use for size bucket (n1, p2)."""
    c1 = C.c_int32(0)
    cr = C.c_int32(0)
    _lib.call("gls_tile_get", C.c_int32(n1), C.c_int32(p2), C.byref(c1), C.byref(cr))
    return c1.value, cr.value


def set_grid_cost(n1, p2, ps_per_cell):
    """Override the cost per grid cell (picoseconds) by which plans made from
    now on choose grid sizes, for bucket (n1, p2); <= 0 removes it."""
    _lib.call("gls_grid_cost_set", C.c_int32(n1), C.c_int32(p2), C.c_double(ps_per_cell))


def get_grid_cost(n1, p2):
    """(cost per cell in picoseconds, source) of bucket (n1, p2); source 0 is
    the model, 1 the measured table, 2 an override."""
    c = C.c_double(0)
    src = C.c_int32(0)
    _lib.call("gls_grid_cost_get", C.c_int32(n1), C.c_int32(p2), C.byref(c), C.byref(src))
    return c.value, src.value


def set_launch_items(items):
    """Cap the work-items of one NFFT-pass launch (at most and by default
    2^31; a size bucket with more runs as several launches over consecutive
    curves); ``items <= 0`` restores the default."""
    _lib.call("gls_launch_items_set", C.c_int64(int(items)))


class Plan:
    """Owns a ``gls_plan`` handle."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def lomb(cls, curves, sigma, m, mode, reg_c, reg_s, input_kind, method,
             grid_chunk_bytes=0, window=0, window_beta=0.0, point_capacity=0, precision=0,
             spread=0):
        _lib.init(0)
        n = len(curves)
        arr = (_lib.gls_curve * max(n, 1))()
        for i, c in enumerate(curves):
            arr[i] = c
        prm = _lib.gls_params(sigma=float(sigma), m=int(m), mode=int(mode),
                              reg_c=float(reg_c), reg_s=float(reg_s),
                              input_kind=int(input_kind), method=int(method),
                              window=int(window), window_beta=float(window_beta),
                              grid_chunk_bytes=int(grid_chunk_bytes),
                              point_capacity=int(point_capacity), precision=int(precision),
                              spread=int(spread))
        h = C.c_void_p()
        _lib.call("gls_lomb_plan_create", C.byref(h), arr, C.c_int32(n), C.byref(prm))
        return cls(h.value)

    def rebind(self, curves):
        """Put other curves (same count and grids) on this plan (gls_plan_rebind)."""
        n = len(curves)
        arr = (_lib.gls_curve * max(n, 1))()
        for i, c in enumerate(curves):
            arr[i] = c
        _lib.call("gls_plan_rebind", self.handle, arr, C.c_int32(n))
        return self

    @classmethod
    def nfft(cls, items, sigma, m):
        _lib.init(0)
        n = len(items)
        arr = (_lib.gls_nfft_item * max(n, 1))()
        for i, c in enumerate(items):
            arr[i] = c
        h = C.c_void_p()
        _lib.call("gls_nfft_plan_create", C.byref(h), arr, C.c_int32(n), C.c_double(sigma),
                  C.c_int32(m))
        return cls(h.value)

    @property
    def workspace_bytes(self):
        v = C.c_size_t(0)
        _lib.call("gls_plan_workspace_bytes", self.handle, C.byref(v))
        return v.value

    @property
    def output_count(self):
        v = C.c_int64(0)
        _lib.call("gls_plan_output_count", self.handle, C.byref(v))
        return v.value

    def grid_size(self, i):
        v = C.c_int64(0)
        _lib.call("gls_plan_grid_size", self.handle, C.c_int32(i), C.byref(v))
        return v.value

    def geometry(self, i):
        """(n, n1, n2, b) of curve i: grid size, four-step split, Gaussian width."""
        n = C.c_int64(0)
        n1 = C.c_int32(0)
        n2 = C.c_int32(0)
        b = C.c_double(0)
        _lib.call("gls_plan_curve_geometry", self.handle, C.c_int32(i), C.byref(n), C.byref(n1),
                  C.byref(n2), C.byref(b))
        return n.value, n1.value, n2.value, b.value

    def tiles(self, i):
        """(c1, cr) of curve i's size bucket: grid columns per pass-1
        workgroup, grid rows per pass-2 workgroup."""
        c1 = C.c_int32(0)
        cr = C.c_int32(0)
        _lib.call("gls_plan_curve_tiles", self.handle, C.c_int32(i), C.byref(c1), C.byref(cr))
        return c1.value, cr.value

    def spread(self, i):
        """SPREAD_FLOAT or SPREAD_FIXED: how pass 1 accumulates curve i's
        window taps (the plan's choice under SPREAD_AUTO)."""
        v = C.c_int32(0)
        _lib.call("gls_plan_curve_spread", self.handle, C.c_int32(i), C.byref(v))
        return v.value

    def set_fork(self, enable=True):
        """Run independent size buckets concurrently on the library's streams."""
        _lib.call("gls_plan_set_fork", self.handle, C.c_int32(1 if enable else 0))
        return self

    def enable_timing(self, enable=True):
        _lib.call("gls_plan_enable_timing", self.handle, C.c_int32(1 if enable else 0))

    STAGE_NAMES = {0: "prep", 1: "pass1", 2: "pass2", 3: "dirsum", 4: "lowfreq"}

    def stage_times(self):
        """[(stage, grid n, ms, algorithmic bytes)] of the last timed execute
        (call after the stream has completed)."""
        n = C.c_int32(0)
        _lib.call("gls_plan_stage_count", self.handle, C.byref(n))
        k = (C.c_int32 * max(n.value, 1))()
        p = (C.c_int32 * max(n.value, 1))()
        ms = (C.c_float * max(n.value, 1))()
        by = (C.c_double * max(n.value, 1))()
        _lib.call("gls_plan_stage_times", self.handle, k, p, ms, by, n)
        return [(self.STAGE_NAMES.get(k[i], str(k[i])), p[i], ms[i], by[i])
                for i in range(n.value)]

    def stage_bucket(self, i):
        """(n1, p2, c1, cr, spread, streamed) of the size bucket stage i runs
        (zeros for stages other than passes 1 and 2; streamed: pass 1 stores
        the grids non-temporally)."""
        v = [C.c_int32(0) for _ in range(6)]
        _lib.call("gls_plan_stage_bucket", self.handle, C.c_int32(i), *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def destroy(self):
        h, self.handle = getattr(self, "handle", None), None
        if h and _lib._lib is not None:
            _lib.call("gls_plan_destroy", C.c_void_p(h))

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # pragma: no cover
            pass


class Graph:
    """Owns a ``gls_graph`` (a captured execute)."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream):
        _lib.call("gls_graph_launch", self.handle, _handle(stream))

    def update(self, plan, stream):
        """Upload a rebound plan's descriptors into the graph's workspace on
        ``stream`` (gls_graph_update); the next launch computes its curves."""
        _lib.call("gls_graph_update", self.handle, plan.handle, _handle(stream))

    def destroy(self):
        h, self.handle = getattr(self, "handle", None), None
        if h and _lib._lib is not None:
            _lib.call("gls_graph_destroy", C.c_void_p(h))

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # pragma: no cover
            pass


class LombBatch:
    """Many light curves, planned and run together.

    Parameters
    ----------
    data : list of (t, y, dy) (``input_kind='raw'``) or (t, yw, w) with
        per-curve ``ybar``/``yy`` (``input_kind='weighted'``)
    freqs : list of linear frequency grids (one per curve); with
        ``custom_freqs=True`` any 1-D arrays of frequencies, evaluated by
        direct sums at exactly those values
    grid_chunk_bytes : if > 0, NFFT grids run in chunks of this many bytes
        that reuse one workspace region (smaller workspace, slightly slower)
    window, window_beta : spreading window of the NFFT path: the reference's
        Gaussian (``WINDOW_GAUSSIAN``) or the exponential of semicircle
        (``WINDOW_ES``, shape ``window_beta``; 0 picks the default)
    precision : ``PRECISION_FP64_COORDS`` (FP64 time offsets, cells and
        phases; the default) or ``PRECISION_FP32`` (the reference's FP32 time
        arithmetic, for comparison)
    spread : ``SPREAD_AUTO`` (the default: fixed point for curves of at
        least one point per 80 grid cells on pass-1 columns of at least 64
        cells, FP32 atomics for the others),
        ``SPREAD_FLOAT`` (FP32 atomic accumulation for every curve) or
        ``SPREAD_FIXED`` (integer fixed-point accumulation for every curve:
        bitwise reproducible results; slower than FP32 atomics on sparse
        curves, faster on dense ones)
    """

    def __init__(self, data, freqs, sigma=2.0, m=8, mode=_lib.MODE_FLOATING_MEAN,
                 amplitude_prior=None, use_fft=True, use_double=False,
                 input_kind="raw", ybar=None, yy=None, custom_freqs=False,
                 grid_chunk_bytes=0, window=_lib.WINDOW_GAUSSIAN, window_beta=0.0,
                 precision=_lib.PRECISION_FP64_COORDS, spread=_lib.SPREAD_AUTO):
        if len(freqs) != len(data):
            raise ValueError("need one frequency grid per light curve")
        self.ncurves = len(data)
        self.use_double = bool(use_double)
        self.input_kind = _lib.INPUT_RAW if input_kind == "raw" else _lib.INPUT_WEIGHTED
        self.out_dtype = np.float64 if use_double else np.float32
        lam = 0.0 if amplitude_prior is None else float(amplitude_prior) ** -2
        curves, obs, out = [], 0, 0
        self.n_obs = np.empty(self.ncurves, dtype=np.int64)
        self.nf = np.empty(self.ncurves, dtype=np.int64)
        self.out_offsets = np.empty(self.ncurves, dtype=np.int64)
        self.custom = bool(custom_freqs)
        for i, (d, fr) in enumerate(zip(data, freqs)):
            if self.custom:
                fr = np.asarray(fr, dtype=np.float64)
                if fr.ndim != 1 or len(fr) < 1 or not np.all(np.isfinite(fr)):
                    raise ValueError("custom frequencies must be a finite 1-D array")
                f0, df, nf = 0.0, 1.0, len(fr)
            else:
                f0, df, nf = linear_grid(fr)
            n = len(d[0])
            c = _lib.gls_curve(obs_offset=obs, n_obs=n, nf=nf, f0=f0, df=df, out_offset=out,
                               ybar=0.0 if ybar is None else float(ybar[i]),
                               yy=1.0 if yy is None else float(yy[i]))
            curves.append(c)
            self.n_obs[i] = n
            self.nf[i] = nf
            self.out_offsets[i] = out
            obs += n
            out += nf
        self.total_obs = obs
        self.total_out = out
        if self.custom:
            method = _lib.METHOD_DIRECT_CUSTOM
            self._freqs = np.concatenate([np.asarray(f, dtype=np.float64) for f in freqs])
        else:
            method = _lib.METHOD_NFFT if use_fft else _lib.METHOD_DIRECT
            self._freqs = None
        self.plan = Plan.lomb(curves, sigma, m, mode, lam, lam, self.input_kind, method,
                              grid_chunk_bytes, window, window_beta, precision=precision,
                              spread=spread)
        # host staging of the inputs
        t = np.empty(obs, dtype=np.float64)
        if self.input_kind == _lib.INPUT_RAW:
            a = np.empty(obs, dtype=np.float64)
            b = np.empty(obs, dtype=np.float64)
        else:
            a = np.empty(obs, dtype=np.float32)
            b = np.empty(obs, dtype=np.float32)
        o = 0
        for d in data:
            n = len(d[0])
            t[o:o + n] = d[0]
            a[o:o + n] = d[1]
            b[o:o + n] = d[2]
            o += n
        self._host = (t, a, b)
        self.d_t = self.d_a = self.d_b = self.d_power = self.d_ws = self.d_freqs = None

    # ---- device side ---------------------------------------------------
    def allocate(self):
        t, a, b = self._host
        self.d_t = DeviceArray(max(len(t), 1), np.float64)
        self.d_a = DeviceArray(max(len(a), 1), a.dtype)
        self.d_b = DeviceArray(max(len(b), 1), b.dtype)
        self.d_power = DeviceArray(max(self.total_out, 1), self.out_dtype)
        self.d_ws = DeviceArray(self.plan.workspace_bytes, np.uint8)
        if self.custom:
            self.d_freqs = DeviceArray(max(self.total_out, 1), np.float64)
        return self

    def upload(self, stream=None):
        if self.d_t is None:
            self.allocate()
        t, a, b = self._host
        if len(t):
            self.d_t.set_async(t, stream=stream)
            self.d_a.set_async(a, stream=stream)
            self.d_b.set_async(b, stream=stream)
        if self.custom:
            self.d_freqs.set_async(self._freqs, stream=stream)
        return self

    def execute(self, stream=None):
        if self.custom:
            _lib.call("gls_lomb_execute_custom", self.plan.handle, self.d_t.ptr, self.d_a.ptr,
                      self.d_b.ptr, self.d_freqs.ptr, self.d_power.ptr,
                      C.c_int32(1 if self.use_double else 0), self.d_ws.ptr, _handle(stream))
            return self
        fn = "gls_lomb_execute_f64" if self.use_double else "gls_lomb_execute"
        _lib.call(fn, self.plan.handle, self.d_t.ptr, self.d_a.ptr, self.d_b.ptr,
                  self.d_power.ptr, self.d_ws.ptr, _handle(stream))
        return self

    def download_async(self, out=None, stream=None):
        if out is None:
            out = pinned_empty(max(self.total_out, 1), self.out_dtype)
        self.d_power.get_async(out, stream=stream, count=self.total_out)
        return out

    def split(self, flat):
        return [flat[o:o + n] for o, n in zip(self.out_offsets, self.nf)]

    def run(self, stream=None):
        """Upload, execute, download; synchronous.  Returns per-curve powers."""
        s = stream or Stream()
        self.upload(stream=s)
        self.execute(stream=s)
        flat = self.download_async(stream=s)
        s.synchronize()
        return [np.array(p) for p in self.split(flat)]

    def capture(self, stream):
        """A HIP graph of this batch's execute on its current device buffers
        (``allocate``/``upload`` first); ``launch(stream)`` reruns it on the
        current contents of the inputs."""
        if self.custom:
            raise NotImplementedError("custom-frequency batches are not captured")
        if self.d_t is None:
            self.allocate()
        g = C.c_void_p()
        _lib.call("gls_plan_graph_create", self.plan.handle, self.d_t.ptr, self.d_a.ptr,
                  self.d_b.ptr, self.d_power.ptr, C.c_int32(1 if self.use_double else 0),
                  self.d_ws.ptr, _handle(stream), C.byref(g))
        return Graph(g.value)

    def topk(self, k, mask=None, stream=None):
        """Top-k (power, index) per curve from the device result."""
        tp = DeviceArray(max(self.ncurves * k, 1), np.float32)
        ti = DeviceArray(max(self.ncurves * k, 1), np.int32)
        dmask = None
        if mask is not None:
            dmask = DeviceArray.to_device(np.ascontiguousarray(mask, dtype=np.uint8))
        if self.use_double:
            raise NotImplementedError("topk on f64 powers")
        _lib.call("gls_topk", self.plan.handle, self.d_ws.ptr, self.d_power.ptr,
                  dmask.ptr if dmask is not None else None, C.c_int32(k), tp.ptr, ti.ptr,
                  _handle(stream))
        # read back on the stream of the top-k launch (the null stream does not
        # wait for a non-blocking stream)
        n = max(self.ncurves * k, 1)
        p = tp.get_async(pinned_empty(n, np.float32), stream=stream, count=n)
        i = ti.get_async(pinned_empty(n, np.int32), stream=stream, count=n)
        if stream is None:
            _lib.call("gls_device_synchronize")
        else:
            _lib.call("gls_stream_synchronize", _handle(stream))
        nk = self.ncurves * k
        return (np.array(p[:nk]).reshape(self.ncurves, k),
                np.array(i[:nk]).reshape(self.ncurves, k))


class NFFTBatch:
    """Adjoint NFFTs sum_j y_j exp(2 pi i (f0 + k df) t_j), k < nf, of many items."""

    def __init__(self, items, sigma=2.0, m=8):
        # items: list of (t, y, f0, df, nf)
        descs, obs, out = [], 0, 0
        self.nf, self.out_offsets = [], []
        for (t, y, f0, df, nf) in items:
            n = len(t)
            descs.append(_lib.gls_nfft_item(obs_offset=obs, n_obs=n, nf=int(nf), f0=float(f0),
                                            df=float(df), out_offset=out))
            self.nf.append(int(nf))
            self.out_offsets.append(out)
            obs += n
            out += int(nf)
        self.total_out = out
        self.plan = Plan.nfft(descs, sigma, m)
        self._t = np.concatenate([np.asarray(i[0], np.float64) for i in items]) if items else \
            np.zeros(0)
        self._y = np.concatenate([np.asarray(i[1], np.float64) for i in items]) if items else \
            np.zeros(0)

    def launch(self, stream):
        """Enqueue upload, execute and download on ``stream``; returns the
        page-locked complex64 result (valid once the stream has completed) and
        the device buffers the enqueued work uses, to be kept alive until then."""
        d_t = DeviceArray.to_device(self._t if len(self._t) else np.zeros(1), stream=stream)
        d_y = DeviceArray.to_device(self._y if len(self._y) else np.zeros(1), stream=stream)
        d_out = DeviceArray(max(2 * self.total_out, 1), np.float32)
        d_ws = DeviceArray(self.plan.workspace_bytes, np.uint8)
        _lib.call("gls_nfft_execute", self.plan.handle, d_t.ptr, d_y.ptr, C.c_int32(0),
                  d_out.ptr, d_ws.ptr, _handle(stream))
        out = pinned_empty(max(2 * self.total_out, 1), np.float32)
        d_out.get_async(out, stream=stream, count=2 * self.total_out)
        return out[: 2 * self.total_out].view(np.complex64), (d_t, d_y, d_out, d_ws, self.plan)

    def split(self, flat):
        return [flat[o:o + n] for o, n in zip(self.out_offsets, self.nf)]

    def run(self, stream=None):
        s = stream or Stream()
        flat, keep = self.launch(s)
        s.synchronize()
        return [np.array(c) for c in self.split(flat)]


def gaussian_grid(t, y, nf, sigma, m, samples_per_peak=1.0, stream=None):
    """Real Gaussian grid in the reference's geometry (just_return_gridded_data)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    ng = int(sigma * nf)
    b = 2.0 * sigma * m / ((2.0 * sigma - 1.0) * np.pi)
    s = stream or Stream()
    d_t = DeviceArray.to_device(t, stream=s)
    d_y = DeviceArray.to_device(y, stream=s)
    d_g = DeviceArray(ng, np.float32)
    _lib.call("gls_gaussian_grid", d_t.ptr, d_y.ptr, C.c_int32(len(t)), C.c_int32(ng),
              C.c_int32(m), C.c_double(b), C.c_double(float(t.min())),
              C.c_double(float(t.max())), C.c_double(samples_per_peak), d_g.ptr, s.handle)
    out = d_g.get_async(stream=s)
    s.synchronize()
    return out


class MultiHarmonicBatch:
    """Multi-harmonic floating-mean GLS (H = 1 .. 4 harmonics) of many curves.

    Per curve the 3H harmonic sums (yw at n f, n <= H; w at n f, n <= 2H) are
    adjoint NFFTs of one NFFT plan; ``gls_mh_power`` then assembles and solves
    the 2H x 2H system per frequency in FP64.  Weights, weighted mean and
    variance are formed on the host in FP64, as LombScargleMemory.setdata does
    (cuvarbase/lombscargle.py:235-293)."""

    def __init__(self, data, freqs, nharmonics=2, sigma=2.0, m=8, amplitude_prior=None,
                 use_double=False):
        if not 1 <= int(nharmonics) <= 4:
            raise ValueError("nharmonics must be in [1, 4]")
        if len(freqs) != len(data):
            raise ValueError("need one frequency grid per light curve")
        if len(data) > 65535:
            raise ValueError("at most 65535 curves per multi-harmonic batch")
        H = self.nharmonics = int(nharmonics)
        self.use_double = bool(use_double)
        self.out_dtype = np.float64 if use_double else np.float32
        lam = 0.0 if amplitude_prior is None else float(amplitude_prior) ** -2
        items, curves, ts, ys = [], [], [], []
        obs = out = pw = 0
        self.nf, self.out_offsets = [], []
        for (t, y, dy), fr in zip(data, freqs):
            f0, df, K = linear_grid(fr)
            t = np.asarray(t, np.float64)
            y = np.asarray(y, np.float64)
            w = 1.0 / np.asarray(dy, np.float64) ** 2
            w /= np.sum(w)
            ybar = float(np.dot(w, y))
            yw = w * (y - ybar)
            yy = float(np.dot(w, (y - ybar) ** 2))
            n = len(t)
            ts += [t, t]
            ys += [yw, w]
            rows = [(obs, n) for _ in range(H)] + [(obs + n, n) for _ in range(2 * H)]
            harm = list(range(1, H + 1)) + list(range(1, 2 * H + 1))
            curves.append(_lib.gls_mh_curve(sums_offset=out, out_offset=pw, nf=K, pad=0, yy=yy,
                                            ybar=float(np.sum(yw)), sum_w=float(np.sum(w)),
                                            lam=lam))
            for (o, cnt), h in zip(rows, harm):
                items.append(_lib.gls_nfft_item(obs_offset=o, n_obs=cnt, nf=K, f0=h * f0,
                                                df=h * df, out_offset=out))
                out += K
            self.nf.append(K)
            self.out_offsets.append(pw)
            obs += 2 * n
            pw += K
        self.total_sums = out
        self.total_out = pw
        self.max_nf = max(self.nf) if self.nf else 0
        self.plan = Plan.nfft(items, sigma, m)
        self._t = np.concatenate(ts) if ts else np.zeros(1)
        self._y = np.concatenate(ys) if ys else np.zeros(1)
        arr = (_lib.gls_mh_curve * max(len(curves), 1))(*curves)
        self._curves = np.frombuffer(arr, dtype=np.uint8).copy()
        self.ncurves = len(curves)

    def run(self, stream=None):
        s = stream or Stream()
        d_t = DeviceArray.to_device(self._t, stream=s)
        d_y = DeviceArray.to_device(self._y, stream=s)
        d_c = DeviceArray.to_device(self._curves, stream=s)
        d_sums = DeviceArray(max(2 * self.total_sums, 1), np.float32)
        d_ws = DeviceArray(self.plan.workspace_bytes, np.uint8)
        d_p = DeviceArray(max(self.total_out, 1), self.out_dtype)
        _lib.call("gls_nfft_execute", self.plan.handle, d_t.ptr, d_y.ptr, C.c_int32(0),
                  d_sums.ptr, d_ws.ptr, _handle(s))
        _lib.call("gls_mh_power", d_sums.ptr, d_c.ptr, C.c_int32(self.ncurves),
                  C.c_int32(self.nharmonics), C.c_int32(self.max_nf), d_p.ptr,
                  C.c_int32(1 if self.use_double else 0), _handle(s))
        flat = pinned_empty(max(self.total_out, 1), self.out_dtype)
        d_p.get_async(flat, stream=s, count=self.total_out)
        s.synchronize()
        return [np.array(flat[o:o + n]) for o, n in zip(self.out_offsets, self.nf)]
