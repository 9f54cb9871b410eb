"""Adjoint NFFT: NFFTMemory, nfft_adjoint_async, NFFTAsyncProcess
(drop-in for cuvarbase/cunfft.py:23-542).

The adjoint NFFT of a curve is ``ghat[k] = sum_j y_j exp(2 pi i f_k t_j)``
with ``f_k = f0 + k df``, ``f0 = minimum_frequency`` and
``df = 1/(samples_per_peak (tmax - tmin))`` -- the quantity the reference's
precompute_psi -> fast_gaussian_grid -> nfft_shift -> cuFFT -> normalize
sequence approximates (cuvarbase/cunfft.py:163-308).  Here it is computed by
the libgls pipeline: FP64 cell coordinates and band-centring modulation, LDS
Gaussian spreading fused into the first FFT pass, deconvolution and the
absolute-time phase fused into the second.  ``t`` stays FP64 on the device
(the reference casts it to FP32, cuvarbase/cunfft.py:151).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._device import DeviceArray, pinned_zeros, synchronize, _handle
from .core import GPUAsyncProcess
from .engine import Plan


def _gauss_b(sigma, m):
    """b = 2 sigma m / ((2 sigma - 1) pi)  (cuvarbase/cunfft.py:56-57)."""
    return float(2 * sigma * m) / ((2 * sigma - 1) * np.pi)


class NFFTMemory(object):
    """Device and pinned host buffers of one adjoint NFFT
    (cuvarbase/cunfft.py:23-160)."""

    def __init__(self, sigma, stream, m, use_double=False, precomp_psi=True, **kwargs):
        self.sigma = sigma
        self.stream = stream
        self.m = m
        self.use_double = use_double
        self.precomp_psi = precomp_psi
        self.real_type = np.float32 if not self.use_double else np.float64
        self.complex_type = np.complex64 if not self.use_double else np.complex128
        self.other_settings = dict(kwargs)
        self.t = kwargs.get('t', None)
        self.y = kwargs.get('y', None)
        self.f0 = kwargs.get('f0', 0.)
        self.n0 = kwargs.get('n0', None)
        self.nf = kwargs.get('nf', None)
        self.t_g = kwargs.get('t_g', None)
        self.y_g = kwargs.get('y_g', None)
        self.ghat_g = kwargs.get('ghat_g', None)
        self.ghat_c = kwargs.get('ghat_c', None)
        self.b = _gauss_b(self.sigma, self.m)
        self.n = None
        self._plans = {}
        self._ws = None

    def allocate_data(self, **kwargs):
        self.n0 = kwargs.get('n0', self.n0)
        self.nf = kwargs.get('nf', self.nf)
        assert self.n0 is not None
        assert self.nf is not None
        self.t_g = DeviceArray(max(self.n0, 1), np.float64, zero=True)
        self.y_g = DeviceArray(max(self.n0, 1), np.float64, zero=True)
        return self

    def allocate_precomp_psi(self, **kwargs):
        # the psi factors are recomputed inside the fused spreading pass
        self.n0 = kwargs.get('n0', self.n0)
        return self

    def allocate_grid(self, **kwargs):
        self.nf = kwargs.get('nf', self.nf)
        assert self.nf is not None
        self.n = int(self.sigma * self.nf)
        # the transform output (nf complex64 values)
        self.ghat_g = DeviceArray(max(2 * self.nf, 1), np.float32, zero=True)
        return self

    def allocate_pinned_cpu(self, **kwargs):
        self.nf = kwargs.get('nf', self.nf)
        assert self.nf is not None
        self.ghat_c = pinned_zeros(self.nf, np.complex64)
        return self

    def is_ready(self):
        assert self.n0 == len(self.t_g) or self.n0 == 0
        assert self.ghat_g is not None

    def allocate(self, **kwargs):
        self.n0 = kwargs.get('n0', self.n0)
        self.nf = kwargs.get('nf', self.nf)
        assert self.n0 is not None
        assert self.nf is not None
        self.n = int(self.sigma * self.nf)
        self.allocate_data(**kwargs)
        self.allocate_grid(**kwargs)
        self.allocate_pinned_cpu(**kwargs)
        return self

    def transfer_data_to_gpu(self, **kwargs):
        t = kwargs.get('t', self.t)
        y = kwargs.get('y', self.y)
        assert t is not None
        assert y is not None
        if len(t):
            self.t_g.set_async(np.asarray(t, np.float64), stream=self.stream)
            self.y_g.set_async(np.asarray(y, np.float64), stream=self.stream)

    def transfer_nfft_to_cpu(self, **kwargs):
        self.ghat_g.get_async(self.ghat_c.view(np.float32), stream=self.stream,
                              count=2 * self.nf)

    def fromdata(self, t, y, allocate=True, **kwargs):
        self.tmin = min(t)
        self.tmax = max(t)
        self.t = np.asarray(t).astype(np.float64)
        self.y = np.asarray(y).astype(self.real_type).astype(np.float64)
        self.n0 = kwargs.get('n0', len(t))
        self.nf = kwargs.get('nf', self.nf)
        if self.nf is not None and allocate:
            self.allocate(**kwargs)
        return self

    # -- plan cache (one plan per (f0, df)) -----------------------------------
    def _plan(self, f0, df):
        key = (float(f0), float(df), int(self.nf), int(self.n0))
        p = self._plans.get(key)
        if p is None:
            item = _lib.gls_nfft_item(obs_offset=0, n_obs=int(self.n0), nf=int(self.nf),
                                      f0=float(f0), df=float(df), out_offset=0)
            p = Plan.nfft([item], self.sigma, self.m)
            self._plans = {key: p}
            need = p.workspace_bytes
            if self._ws is None or self._ws.nbytes < need:
                self._ws = DeviceArray(need, np.uint8)
        return p


def nfft_adjoint_async(memory, functions, minimum_frequency=0., block_size=256,
                       just_return_gridded_data=False, use_grid=None, fast_grid=True,
                       transfer_to_device=True, transfer_to_host=True, precomp_psi=True,
                       samples_per_peak=1, **kwargs):
    """Asynchronous adjoint NFFT of one curve (cuvarbase/cunfft.py:163-308).

    ``functions`` is the 5-tuple of NFFTAsyncProcess.function_tuple; the
    five reference stages are fused here, so the tuple is accepted for
    signature compatibility.  ``fast_grid`` selects nothing (both reference
    gridding variants compute the same grid).  ``use_grid`` (grid injection for
    debugging) is not supported and raises NotImplementedError.
    """
    if use_grid is not None:
        raise NotImplementedError("use_grid: the spread grid is never materialised in "
                                  "the fused pipeline")
    stream = memory.stream
    if transfer_to_device:
        memory.transfer_data_to_gpu()
    tmin, tmax = float(memory.tmin), float(memory.tmax)
    if just_return_gridded_data:
        ng = int(memory.sigma * memory.nf)
        grid = DeviceArray(ng, np.float32)
        _lib.call("gls_gaussian_grid", memory.t_g.ptr, memory.y_g.ptr, C.c_int32(memory.n0),
                  C.c_int32(ng), C.c_int32(memory.m), C.c_double(memory.b),
                  C.c_double(tmin), C.c_double(tmax), C.c_double(samples_per_peak),
                  grid.ptr, _handle(stream))
        out = grid.get_async(stream=stream)
        if stream is not None:
            stream.synchronize()
        else:
            synchronize()
        return out
    df = 1.0 / (samples_per_peak * (tmax - tmin))
    plan = memory._plan(minimum_frequency, df)
    _lib.call("gls_nfft_execute", plan.handle, memory.t_g.ptr, memory.y_g.ptr, C.c_int32(0),
              memory.ghat_g.ptr, memory._ws.ptr, _handle(stream))
    if transfer_to_host:
        memory.transfer_nfft_to_cpu()
    return memory.ghat_c


class _Fused(object):
    """Placeholder bound in function_tuple for a reference stage that the fused
    pipeline performs inside another launch."""

    def __init__(self, name, inside):
        self.name = name
        self.inside = inside

    def __call__(self, *args, **kwargs):
        raise NotImplementedError(f"{self.name} is fused into {self.inside} in libgls")

    def __repr__(self):
        return f"<fused stage {self.name} (in {self.inside})>"


class NFFTAsyncProcess(GPUAsyncProcess):
    """GPUAsyncProcess for the adjoint NFFT (cuvarbase/cunfft.py:311-542).

    Parameters as the reference: ``sigma`` (default 4), ``m`` (8),
    ``autoset_m`` (False), ``tol`` (1e-8), ``block_size``, ``use_double``,
    ``use_fast_math`` (accepted; the HIP build has no fast-math switch).
    """

    def __init__(self, *args, **kwargs):
        super(NFFTAsyncProcess, self).__init__(*args, **kwargs)
        self.sigma = kwargs.get('sigma', 4)
        self.m = kwargs.get('m', 8)
        self.autoset_m = kwargs.get('autoset_m', False)
        self.block_size = kwargs.get('block_size', 256)
        self.use_double = kwargs.get('use_double', False)
        self.m_tol = kwargs.get('tol', 1E-8)
        self.module_options = []
        if kwargs.get('use_fast_math', True):
            self.module_options.append('--use_fast_math')
        self.real_type = np.float64 if self.use_double else np.float32
        self.complex_type = np.complex128 if self.use_double else np.complex64
        self._cpp_defs = dict(BLOCK_SIZE=self.block_size)
        if self.use_double:
            self._cpp_defs['DOUBLE_PRECISION'] = None
        self.function_names = ['precompute_psi', 'fast_gaussian_grid', 'slow_gaussian_grid',
                               'nfft_shift', 'normalize']
        self.allocated_memory = []

    def m_from_C(self, C_, sigma):
        """m = ceil(-ln(C/4) / (pi (1 - 1/(2 sigma - 1)))) (cuvarbase/cunfft.py:378-386)."""
        D = (np.pi * (1. - 1. / (2. * sigma - 1.)))
        return int(np.ceil(-np.log(0.25 * C_) / D))

    def estimate_m(self, N):
        return self.m_from_C(self.m_tol / N, self.sigma)

    def get_m(self, N=None):
        return self.estimate_m(N) if self.autoset_m else self.m

    def _compile_and_prepare_functions(self, **kwargs):
        _lib.load()
        self._select_device()
        self.prepared_functions = {
            'precompute_psi': _Fused('precompute_psi', 'the pass-1 spreading'),
            'fast_gaussian_grid': 'gls_gaussian_grid',
            'slow_gaussian_grid': 'gls_gaussian_grid',
            'nfft_shift': _Fused('nfft_shift', 'the prep band-centring modulation'),
            'normalize': _Fused('normalize', 'pass 2'),
        }
        self.function_tuple = tuple(self.prepared_functions[f] for f in self.function_names)

    def allocate(self, data, **kwargs):
        allocated_memory = []
        if len(data) > len(self.streams):
            self._create_streams(len(data) - len(self.streams))
        for i, (t, y, nf) in enumerate(data):
            m = self.get_m(nf)
            mem = NFFTMemory(self.sigma, self.streams[i], m, use_double=self.use_double,
                             **kwargs)
            allocated_memory.append(mem.fromdata(t, y, nf=nf, allocate=True, **kwargs))
        return allocated_memory

    def run(self, data, memory=None, **kwargs):
        """Adjoint NFFTs of [(t, y, nf), ...]; returns the pinned ``ghat_c``
        arrays, valid after finish() (cuvarbase/cunfft.py:507-542)."""
        if not all(f in self.prepared_functions for f in self.function_names):
            self._compile_and_prepare_functions(**kwargs)
        if memory is None:
            memory = self.allocate(data, **kwargs)
        nfft_kwargs = dict(block_size=self.block_size)
        nfft_kwargs.update(kwargs)
        self._inflight.append(memory)
        return [nfft_adjoint_async(mem, self.function_tuple, **nfft_kwargs) for mem in memory]
