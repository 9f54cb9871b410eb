"""Adjoint NFFT with the Python API of cuvarbase/cunfft.py (NFFTMemory,
nfft_adjoint_async, NFFTAsyncProcess).

The adjoint NFFT of a curve is ``ghat[k] = sum_j y_j exp(2 pi i f_k t_j)``
with ``f_k = f0 + k df``, ``f0 = minimum_frequency`` and
``df = 1/(samples_per_peak (tmax - tmin))`` -- the quantity the reference's
precompute_psi -> fast_gaussian_grid -> nfft_shift -> cuFFT -> normalize
sequence approximates (cuvarbase/cunfft.py:163-308).  libgls computes it with
FP64 cell coordinates and band-centring modulation, Gaussian spreading in LDS
fused into the first pass of a four-step FFT, and deconvolution plus the
absolute-time phase fused into the second.  ``t`` stays FP64 on the device
(the reference casts it to FP32, cuvarbase/cunfft.py:151).

``NFFTAsyncProcess.run(data)`` without preallocated memory runs every item of
``data`` in one planned launch sequence.  With ``use_double=True`` the
results are complex128 arrays, filled by ``finish()`` from the FP32 pipeline
(relative accuracy ~1e-6 of sum |y|, not an FP64 transform).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._device import DeviceArray, pinned_zeros, synchronize, _handle
from .core import GPUAsyncProcess
from .engine import NFFTBatch, Plan


def gaussian_b(sigma, m):
    """Gaussian window width b = 2 sigma m / ((2 sigma - 1) pi) in cells^2
    (cuvarbase/cunfft.py:56-57)."""
    return float(2 * sigma * m) / ((2 * sigma - 1) * np.pi)


def _sync(stream):
    if stream is None:
        synchronize()
    else:
        stream.synchronize()


# attributes of NFFTMemory that keyword arguments may preset, and their defaults
MEMORY_FIELDS = (("t", None), ("y", None), ("f0", 0.0), ("n0", None), ("nf", None),
                 ("t_g", None), ("y_g", None), ("ghat_g", None), ("ghat_c", None))


class NFFTMemory(object):
    """Buffers of one adjoint NFFT.

    ``t_g``, ``y_g``: the points on the device (FP64); ``ghat_g``: the nf
    transform values (complex64) on the device; ``ghat_c``: their host copy,
    page-locked complex64, or complex128 with ``use_double`` (then written by
    ``complete()`` once the stream is done).  The oversampled grid lives in
    the workspace of a one-item plan, made at the first transform and kept for
    the same (f0, df, nf, n0); ``n`` is that plan's grid size.
    """

    def __init__(self, sigma, stream, m, use_double=False, precomp_psi=True, **kwargs):
        self.sigma = sigma
        self.stream = stream
        self.m = m
        self.use_double = bool(use_double)
        self.precomp_psi = precomp_psi
        self.real_type = np.float64 if self.use_double else np.float32
        self.complex_type = np.complex128 if self.use_double else np.complex64
        self.other_settings = dict(kwargs)
        for name, default in MEMORY_FIELDS:
            setattr(self, name, kwargs.get(name, default))
        self.b = gaussian_b(sigma, m)
        self.n = None
        self.tmin = self.tmax = None
        self._plan_key, self._plan, self._ws = None, None, None
        self._staging = None   # page-locked complex64 copy behind a complex128 ghat_c

    # -- allocation ---------------------------------------------------------
    def _sizes(self, kwargs):
        self.n0 = kwargs.get("n0", self.n0)
        self.nf = kwargs.get("nf", self.nf)
        if self.n0 is None or self.nf is None:
            raise AssertionError("NFFTMemory needs n0 and nf before allocation")
        return int(self.n0), int(self.nf)

    def allocate_data(self, **kwargs):
        n0, _ = self._sizes(kwargs)
        self.t_g = DeviceArray(max(n0, 1), np.float64, zero=True)
        self.y_g = DeviceArray(max(n0, 1), np.float64, zero=True)
        return self

    def allocate_precomp_psi(self, **kwargs):
        """Nothing to allocate: the window weights are formed inside the
        spreading pass."""
        self.n0 = kwargs.get("n0", self.n0)
        return self

    def allocate_grid(self, **kwargs):
        """The device array of the nf transform values (the grid itself is
        plan workspace)."""
        _, nf = self._sizes(kwargs)
        if self.n is None:
            self.n = int(self.sigma * nf)
        self.ghat_g = DeviceArray(max(2 * nf, 1), np.float32, zero=True)
        return self

    def allocate_pinned_cpu(self, **kwargs):
        _, nf = self._sizes(kwargs)
        if self.use_double:
            self._staging = pinned_zeros(max(nf, 1), np.complex64)
            self.ghat_c = np.zeros(nf, dtype=np.complex128)
        else:
            self.ghat_c = pinned_zeros(max(nf, 1), np.complex64)[:nf]
        return self

    def is_ready(self):
        assert self.t_g is not None and self.y_g is not None and self.ghat_g is not None
        assert self.n0 is not None and len(self.t_g) >= self.n0
        if self.ghat_c is not None:
            assert len(self.ghat_c) == self.nf

    def allocate(self, **kwargs):
        self._sizes(kwargs)
        self.allocate_data(**kwargs)
        self.allocate_grid(**kwargs)
        self.allocate_pinned_cpu(**kwargs)
        self.allocate_precomp_psi(**kwargs)
        return self

    # -- data ------------------------------------------------------------------
    def fromdata(self, t, y, allocate=True, **kwargs):
        """Keep (t, y) in FP64 and their time range; allocate when nf is known."""
        self.t = np.asarray(t, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.tmin, self.tmax = float(np.min(self.t)), float(np.max(self.t))
        self.n0 = kwargs.get("n0", len(self.t))
        self.nf = kwargs.get("nf", self.nf)
        if allocate and self.nf is not None:
            self.allocate(**kwargs)
        return self

    def transfer_data_to_gpu(self, **kwargs):
        t = kwargs.get("t", self.t)
        y = kwargs.get("y", self.y)
        assert t is not None and y is not None
        if len(t):
            self.t_g.set_async(np.asarray(t, np.float64), stream=self.stream)
            self.y_g.set_async(np.asarray(y, np.float64), stream=self.stream)

    def transfer_nfft_to_cpu(self, **kwargs):
        dst = self._staging if self.use_double else self.ghat_c
        self.ghat_g.get_async(dst.view(np.float32), stream=self.stream, count=2 * self.nf)

    def complete(self):
        """With use_double: widen the finished transfer into ghat_c."""
        if self.use_double and self._staging is not None:
            self.ghat_c[:] = self._staging[: self.nf]

    # -- plan ------------------------------------------------------------------
    def plan_for(self, f0, df):
        """The one-item plan of this memory's points and nf at (f0, df),
        with a workspace large enough for it."""
        key = (float(f0), float(df), int(self.nf), int(self.n0))
        if key != self._plan_key:
            item = _lib.gls_nfft_item(obs_offset=0, n_obs=int(self.n0), nf=int(self.nf),
                                      f0=float(f0), df=float(df), out_offset=0)
            self._plan = Plan.nfft([item], self.sigma, self.m)
            self._plan_key = key
            self.n = self._plan.grid_size(0)
            need = self._plan.workspace_bytes
            if self._ws is None or self._ws.nbytes < need:
                self._ws = DeviceArray(need, np.uint8)
        return self._plan


def nfft_adjoint_async(memory, functions, minimum_frequency=0., block_size=256,
                       just_return_gridded_data=False, use_grid=None, fast_grid=True,
                       transfer_to_device=True, transfer_to_host=True, precomp_psi=True,
                       samples_per_peak=1, defer_upcast=False, **kwargs):
    """Adjoint NFFT of one curve on ``memory.stream`` (cuvarbase/cunfft.py:163-308).

    Returns ``memory.ghat_c``, valid once the stream has completed.
    ``functions`` is NFFTAsyncProcess.function_tuple; the reference's five
    stages run fused in libgls, so it is accepted for the signature only.
    ``fast_grid`` selects nothing (both reference gridding variants compute
    the same grid).  ``just_return_gridded_data`` returns the real Gaussian
    grid in the reference's one-sided geometry (synchronously).  ``use_grid``
    (injecting a grid) is not supported: the spread grid is never
    materialised.  With ``use_double`` the complex128 result is written here
    after the stream completes, unless ``defer_upcast`` (a process's
    ``finish()`` then writes it)."""
    if use_grid is not None:
        raise NotImplementedError("use_grid: the spread grid is never materialised in "
                                  "the fused pipeline")
    stream = memory.stream
    if transfer_to_device:
        memory.transfer_data_to_gpu()
    if just_return_gridded_data:
        ng = int(memory.sigma * memory.nf)
        grid = DeviceArray(ng, np.float32)
        _lib.call("gls_gaussian_grid", memory.t_g.ptr, memory.y_g.ptr, C.c_int32(memory.n0),
                  C.c_int32(ng), C.c_int32(memory.m), C.c_double(memory.b),
                  C.c_double(memory.tmin), C.c_double(memory.tmax),
                  C.c_double(samples_per_peak), grid.ptr, _handle(stream))
        out = grid.get_async(stream=stream)
        _sync(stream)
        return out
    df = 1.0 / (samples_per_peak * (memory.tmax - memory.tmin))
    plan = memory.plan_for(minimum_frequency, df)
    _lib.call("gls_nfft_execute", plan.handle, memory.t_g.ptr, memory.y_g.ptr, C.c_int32(0),
              memory.ghat_g.ptr, memory._ws.ptr, _handle(stream))
    if transfer_to_host:
        memory.transfer_nfft_to_cpu()
        if memory.use_double and not defer_upcast:
            _sync(stream)
            memory.complete()
    return memory.ghat_c


class GriddingLauncher(object):
    """The ``fast_gaussian_grid`` / ``slow_gaussian_grid`` entry of
    NFFTAsyncProcess.function_tuple: ``launcher(memory, samples_per_peak=1)``
    returns the real Gaussian grid of an NFFTMemory (``gls_gaussian_grid``;
    both reference variants compute the same grid)."""

    def __init__(self, name):
        self.name = name

    def __call__(self, memory, samples_per_peak=1, **kwargs):
        return nfft_adjoint_async(memory, None, just_return_gridded_data=True,
                                  samples_per_peak=samples_per_peak, transfer_to_device=False)

    def __repr__(self):
        return f"<libgls gls_gaussian_grid as {self.name}>"


# keyword arguments of NFFTAsyncProcess and their defaults
# (cuvarbase/cunfft.py:349-376)
NFFT_KEYWORDS = (("sigma", 4), ("m", 8), ("autoset_m", False), ("block_size", 256),
                 ("use_double", False))


class NFFTAsyncProcess(GPUAsyncProcess):
    """Adjoint NFFTs of [(t, y, nf), ...] (cuvarbase/cunfft.py:311-542).

    Keyword arguments: ``sigma`` (default 4), ``m`` (8), ``autoset_m``
    (False), ``tol`` (1e-8, for autoset_m), ``block_size``, ``use_double``,
    ``use_fast_math`` (accepted; the HIP build has no fast-math switch), and
    those of GPUAsyncProcess.

    ``function_tuple`` follows ``function_names``: precompute_psi, nfft_shift
    and normalize are ``None`` (they run fused inside the libgls launches:
    the window weights inside the spreading pass, the band shift inside the
    point preparation, deconvolution and phase inside the second FFT pass);
    fast_gaussian_grid and slow_gaussian_grid are GriddingLauncher callables.
    """

    function_names = ["precompute_psi", "fast_gaussian_grid", "slow_gaussian_grid",
                      "nfft_shift", "normalize"]

    def __init__(self, *args, **kwargs):
        super(NFFTAsyncProcess, self).__init__(*args, **kwargs)
        for name, default in NFFT_KEYWORDS:
            setattr(self, name, kwargs.get(name, default))
        self.m_tol = kwargs.get("tol", 1E-8)
        self.module_options = ["--use_fast_math"] if kwargs.get("use_fast_math", True) else []
        self.real_type = np.float64 if self.use_double else np.float32
        self.complex_type = np.complex128 if self.use_double else np.complex64
        self._cpp_defs = dict(BLOCK_SIZE=self.block_size)
        if self.use_double:
            self._cpp_defs["DOUBLE_PRECISION"] = None
        self.allocated_memory = []

    # -- window half-width (cuvarbase/cunfft.py:378-431) ------------------------
    def m_from_C(self, C_, sigma):
        """Smallest m with alias error 4 exp(-m pi (1 - 1/(2 sigma - 1))) below C."""
        return int(np.ceil(-np.log(0.25 * C_) / (np.pi * (1. - 1. / (2. * sigma - 1.)))))

    def estimate_m(self, N):
        return self.m_from_C(self.m_tol / N, self.sigma)

    def get_m(self, N=None):
        return self.estimate_m(N) if self.autoset_m else self.m

    def _compile_and_prepare_functions(self, **kwargs):
        _lib.load()
        self._select_device()
        grid = GriddingLauncher("fast_gaussian_grid")
        self.prepared_functions = dict(precompute_psi=None, fast_gaussian_grid=grid,
                                       slow_gaussian_grid=GriddingLauncher("slow_gaussian_grid"),
                                       nfft_shift=None, normalize=None)
        self.function_tuple = tuple(self.prepared_functions[f] for f in self.function_names)

    def _ready(self, **kwargs):
        if not all(f in self.prepared_functions for f in self.function_names):
            self._compile_and_prepare_functions(**kwargs)

    def allocate(self, data, **kwargs):
        """An NFFTMemory per (t, y, nf) item, item i on stream i."""
        if data:
            self._stream(len(data) - 1)
        self.allocated_memory = [
            NFFTMemory(self.sigma, self.streams[i], self.get_m(nf), use_double=self.use_double,
                       **kwargs).fromdata(t, y, nf=nf, allocate=True, **kwargs)
            for i, (t, y, nf) in enumerate(data)]
        return self.allocated_memory

    def run(self, data, memory=None, minimum_frequency=0., samples_per_peak=1, **kwargs):
        """Adjoint NFFTs of [(t, y, nf), ...] with f_k = minimum_frequency + k df,
        df = 1/(samples_per_peak (max t - min t)) per item; results (nf
        complex values each) are valid after ``finish()``.

        Without ``memory`` all items run as one batch on stream 0; with a list
        of NFFTMemory (from ``allocate``) each item runs on its memory's
        stream through ``nfft_adjoint_async``."""
        self._ready(**kwargs)
        if memory is not None or kwargs.get("just_return_gridded_data", False):
            if memory is None:
                memory = self.allocate(data, **kwargs)
            nfft_kwargs = dict(block_size=self.block_size, minimum_frequency=minimum_frequency,
                               samples_per_peak=samples_per_peak, defer_upcast=True)
            nfft_kwargs.update(kwargs)
            out = [nfft_adjoint_async(mem, self.function_tuple, **nfft_kwargs) for mem in memory]
            self._keep_alive.append(memory)
            self._on_finish.extend(mem.complete for mem in memory if mem.use_double)
            return out
        items = []
        for (t, y, nf) in data:
            t = np.asarray(t, dtype=np.float64)
            df = 1.0 / (samples_per_peak * (np.max(t) - np.min(t)))
            items.append((t, y, minimum_frequency, df, int(nf)))
        m = max((self.get_m(it[4]) for it in items), default=self.m)
        batch = NFFTBatch(items, sigma=self.sigma, m=m)
        flat, keep = batch.launch(self._stream(0))
        self._keep_alive.append(keep)
        parts = batch.split(flat)
        if not self.use_double:
            return parts
        wide = [np.zeros(len(p), dtype=np.complex128) for p in parts]

        def widen():
            for dst, src in zip(wide, parts):
                dst[:] = src
        self._on_finish.append(widen)
        return wide
