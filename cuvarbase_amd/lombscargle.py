"""Generalized Lomb-Scargle periodogram on MI355X with the Python API of
cuvarbase/lombscargle.py.

Public surface kept from the reference: ``get_k0``, ``check_k0``,
``LombScargleMemory``, ``lomb_scargle_async``, ``LombScargleAsyncProcess``
(``run``, ``allocate``, ``allocate_for_single_lc``, ``preallocate``,
``batched_run``, ``batched_run_const_nfreq``, ``finish``, ``autofrequency``,
``memory_requirement``), the FP64 host functions ``mhdirect_sums``,
``add_regularization``, ``mhgls_params_from_sums``, ``mhgls_from_sums``,
``lomb_scargle_direct_sums``, ``fap_baluev`` and ``lomb_scargle_simple``.

Differences from the reference, by design:

* The periodogram is evaluated at the exact frequencies ``f0 + k df`` of the
  grid passed in; the reference evaluates the 2f sums at ``(2k + k0) df`` with
  ``k0 = round(f0/df)`` (cuvarbase/kernels/lomb.cu:206).  ``check_k0`` is still
  asserted so invalid grids fail the same way.
* ``run(data)`` without preallocated memory plans all curves of ``data`` as
  one batch (one launch sequence) instead of one stream per curve.
* Times stay FP64 on the device; spreading/FFT are FP32, phases and the GLS
  determinants FP64.  ``use_double=True`` returns float64 powers from the same
  pipeline.
* The NFFT spreading window is the reference's Gaussian by default;
  ``nfft_window="es"`` (with ``window_beta``, 0 for the default shape) selects
  the exponential-of-semicircle window, which meets |dP| <= 1e-4 in FP32 at
  sigma = 1.35, m = 5 (the batched engine's tuned configuration).
* ``nharmonics`` 2-4 runs the multi-harmonic GLS (the reference raises).
* ``memory_requirement`` reports the device bytes of a real plan (the
  reference raises UnboundLocalError, cuvarbase/lombscargle.py:765); the
  direct-sum path copies results back under ``transfer_to_host`` (the
  reference tests ``transfer_to_device``, cuvarbase/lombscargle.py:657).
"""
from __future__ import annotations

import numpy as np
from scipy.special import gammaln

from . import _lib
from ._device import DeviceArray, pinned_zeros, _handle
from .core import GPUAsyncProcess
from .cunfft import NFFTAsyncProcess, NFFTMemory
from .engine import LombBatch, Plan
from .utils import weights, normalize_light_curves
from .utils import autofrequency as utils_autofreq


def get_k0(freqs):
    """Multiple of the grid spacing nearest to the first frequency, at least 1
    (cuvarbase/lombscargle.py:25-26)."""
    spacing = freqs[1] - freqs[0]
    return max(1, int(round(freqs[0] / spacing)))


def check_k0(freqs, k0=None, rtol=1E-2, atol=1E-7):
    """AssertionError unless the first frequency is k0 grid spacings, within
    rtol spacings + atol (cuvarbase/lombscargle.py:29-33)."""
    spacing = freqs[1] - freqs[0]
    k0 = get_k0(freqs) if k0 is None else k0
    assert abs(k0 * spacing - freqs[0]) < rtol * spacing + atol


def _mode(floating_mean=True, window=False):
    """libgls mode: the spectral window, else floating mean or standard
    (cuvarbase/lombscargle.py:68-70)."""
    if window:
        return _lib.MODE_WINDOW
    return _lib.MODE_FLOATING_MEAN if floating_mean else _lib.MODE_STANDARD


def _prior_lambdas(amplitude_prior, nharmonics):
    """Per-harmonic ridge lambda = prior^-2 (zeros without a prior)."""
    if amplitude_prior is None:
        return np.zeros(nharmonics)
    lam = np.power(np.asarray(amplitude_prior, dtype=np.float64), -2.0)
    return np.broadcast_to(lam, (nharmonics,)).astype(np.float64)


NFFT_WINDOWS = {"gaussian": _lib.WINDOW_GAUSSIAN, "es": _lib.WINDOW_ES}


def _nfft_window(name):
    try:
        return NFFT_WINDOWS[str(name).lower()]
    except KeyError:
        raise ValueError(f"nfft_window must be one of {sorted(NFFT_WINDOWS)}") from None


# settings of LombScargleMemory taken from keyword arguments, with defaults
# (cuvarbase/lombscargle.py:50-127)
MEMORY_SETTINGS = (("k0", 0), ("precomp_psi", True), ("amplitude_prior", None),
                   ("window", False), ("nharmonics", 1), ("use_fft", True),
                   ("floating_mean", True), ("use_double", False), ("n0", None), ("nf", None),
                   ("buffered_transfer", False), ("n0_buffer", None),
                   ("nfft_window", "gaussian"), ("window_beta", 0.0))
# arrays that keyword arguments may hand in
MEMORY_ARRAYS = ("t_g", "yw_g", "w_g", "lsp_g", "lsp_c", "t", "yw", "w")


class LombScargleMemory(object):
    """Device and host buffers of one light curve.

    Device: ``t_g`` (FP64 times), ``yw_g`` and ``w_g`` (FP32 weighted
    residuals and weights) of capacity ``n0`` (``n0_buffer`` with buffered
    transfers); ``lsp_g``, the ``nf`` powers (FP32, FP64 with use_double).
    Host: ``lsp_c``, page-locked powers; with buffered transfers, ``t``,
    ``yw``, ``w`` staging of the same capacity.  The three NFFT grids are
    the workspace of a one-curve plan made at the first launch and kept while
    the frequency grid, size and settings stay the same; ``nfft_mem_yw`` and
    ``nfft_mem_w`` describe the transforms (nf, grid size n).
    """

    def __init__(self, sigma, stream, m, **kwargs):
        self.sigma = sigma
        self.stream = stream
        self.m = m
        for name, default in MEMORY_SETTINGS:
            setattr(self, name, kwargs.get(name, default))
        for name in MEMORY_ARRAYS:
            setattr(self, name, kwargs.get(name, None))
        self.other_settings = dict(kwargs)
        self.mode = _mode(self.floating_mean, self.window)
        self.real_type = np.float64 if self.use_double else np.float32
        self.complex_type = np.complex128 if self.use_double else np.complex64
        if self.use_fft:
            self.nfft_mem_yw = kwargs.get("nfft_mem_yw") or NFFTMemory(
                sigma, stream, m, use_double=self.use_double)
            self.nfft_mem_w = kwargs.get("nfft_mem_w") or NFFTMemory(
                sigma, stream, m, use_double=self.use_double)
        # ridge terms on the diagonal: (lambda_1, lambda_1, ..., lambda_H, lambda_H, 0)
        lam = _prior_lambdas(self.amplitude_prior, self.nharmonics)
        self.reg = np.zeros(2 * self.nharmonics + 1, dtype=self.real_type)
        self.reg[: 2 * self.nharmonics] = np.repeat(lam, 2)
        self.ybar, self.yy = 0., 1.
        self.tmin = self.tmax = None
        self._plan_key, self._plan, self._ws = None, None, None

    def _capacity(self, kwargs):
        """Points the buffers hold: n0, or n0_buffer with buffered transfers."""
        cap = kwargs.get("n0_buffer", self.n0_buffer) if self.buffered_transfer \
            else kwargs.get("n0", self.n0)
        assert cap is not None, "LombScargleMemory needs n0 (or n0_buffer) to allocate"
        return int(cap)

    def _descriptors(self):
        return (self.nfft_mem_yw, self.nfft_mem_w) if self.use_fft else ()

    # -- allocation (cuvarbase/lombscargle.py:129-233) ----------------------------
    def allocate_data(self, **kwargs):
        cap = self._capacity(kwargs)
        self.t_g = DeviceArray(max(cap, 1), np.float64, zero=True)
        self.yw_g = DeviceArray(max(cap, 1), np.float32, zero=True)
        self.w_g = DeviceArray(max(cap, 1), np.float32, zero=True)
        for desc, values in zip(self._descriptors(), (self.yw_g, self.w_g)):
            desc.t_g, desc.y_g, desc.n0 = self.t_g, values, cap
        return self

    def allocate_grids(self, **kwargs):
        self.nf = kwargs.get("nf", self.nf)
        assert self.nf is not None, "LombScargleMemory needs nf to allocate"
        for desc in self._descriptors():
            desc.nf = int(self.nf)
        self.lsp_g = DeviceArray(max(int(self.nf), 1), self.real_type, zero=True)
        return self

    def allocate_pinned_cpu(self, **kwargs):
        nf = kwargs.get("nf", self.nf)
        assert nf is not None, "LombScargleMemory needs nf to allocate"
        self.lsp_c = pinned_zeros(int(nf), self.real_type)
        return self

    def is_ready(self):
        raise NotImplementedError()

    def allocate_buffered_data_arrays(self, **kwargs):
        cap = self._capacity(kwargs)
        self.t = np.zeros(cap, dtype=np.float64)
        self.yw = np.zeros(cap, dtype=np.float32)
        self.w = np.zeros(cap, dtype=np.float32)
        return self

    def allocate(self, **kwargs):
        self.allocate_data(**kwargs)
        self.allocate_grids(**kwargs)
        self.allocate_pinned_cpu(**kwargs)
        if self.buffered_transfer:
            self.allocate_buffered_data_arrays(**kwargs)
        return self

    # -- data (cuvarbase/lombscargle.py:235-334) -----------------------------------
    def setdata(self, **kwargs):
        """Weights, weighted mean and variance in FP64 on the host; yw and w
        are then kept as FP32, t as FP64.  Takes (t, y, dy), or (t, yw, w)
        with ``yy``."""
        t = np.asarray(kwargs.get("t", self.t), dtype=np.float64)
        n0 = int(kwargs.get("n0", len(t)))
        if kwargs.get("dy") is not None:
            assert "w" not in kwargs
            w = weights(kwargs["dy"])
        else:
            w = np.asarray(kwargs.get("w", self.w), dtype=np.float64)
        if kwargs.get("y") is not None:
            assert "yw" not in kwargs
            resid = np.asarray(kwargs["y"], dtype=np.float64)
            self.ybar = float(np.dot(w, resid))
            resid = resid - self.ybar
            yw = w * resid
            self.yy = float(np.dot(w, resid * resid))
        else:
            yw = kwargs.get("yw", self.yw)
            self.ybar, self.yy = 0., kwargs.get("yy", 1.)
        yw32, w32 = np.asarray(yw, np.float32), np.asarray(w, np.float32)
        if self.buffered_transfer:
            if self.t is None or self.yw is None or self.w is None:
                self.allocate_buffered_data_arrays(**kwargs)
            assert n0 <= len(self.t), "more points than the buffer holds"
            self.t[:n0], self.yw[:n0], self.w[:n0] = t[:n0], yw32[:n0], w32[:n0]
        else:
            self.t, self.yw, self.w = t, yw32, w32
        self.n0 = n0
        self.tmin = float(np.min(t[:n0])) if n0 else 0.0
        self.tmax = float(np.max(t[:n0])) if n0 else 0.0
        for desc in self._descriptors():
            desc.tmin, desc.tmax, desc.n0 = self.tmin, self.tmax, n0
        return self

    def transfer_data_to_gpu(self, **kwargs):
        assert self.t is not None and self.yw is not None and self.w is not None
        if self.n0:
            for dev, host in ((self.t_g, self.t), (self.yw_g, self.yw), (self.w_g, self.w)):
                dev.set_async(host[: self.n0], stream=self.stream)

    def transfer_lsp_to_cpu(self, **kwargs):
        self.lsp_g.get_async(self.lsp_c, stream=self.stream, count=self.nf)

    def fromdata(self, **kwargs):
        self.setdata(**kwargs)
        if kwargs.get("allocate", True):
            self.allocate(**kwargs)
        return self

    def set_gpu_arrays_to_zero(self, **kwargs):
        for dev in (self.t_g, self.yw_g, self.w_g, self.lsp_g):
            if dev is not None:
                dev.fill(0, stream=self.stream)
        if self.buffered_transfer:
            for host in (self.t, self.yw, self.w):
                if host is not None:
                    host[:] = 0

    # -- plan ------------------------------------------------------------------
    def plan_for(self, freqs, method):
        """One-curve plan (weighted input) of this memory's points at the
        linear grid ``freqs``, with a workspace large enough for it."""
        f0, df, nf = float(freqs[0]), float(freqs[1] - freqs[0]), len(freqs)
        win = _nfft_window(self.nfft_window)
        key = (f0, df, nf, int(self.n0), method, self.mode, float(self.reg[0]),
               float(self.reg[1]), float(self.ybar), float(self.yy), win,
               float(self.window_beta))
        if key != self._plan_key:
            curve = _lib.gls_curve(obs_offset=0, n_obs=int(self.n0), nf=nf, f0=f0, df=df,
                                   out_offset=0, ybar=float(self.ybar), yy=float(self.yy))
            self._plan = Plan.lomb([curve], self.sigma, self.m, self.mode, float(self.reg[0]),
                                   float(self.reg[1]), _lib.INPUT_WEIGHTED, method,
                                   window=win, window_beta=float(self.window_beta),
                                   spread=_lib.SPREAD_AUTO)
            self._plan_key = key
            if method == _lib.METHOD_NFFT:
                for desc in self._descriptors():
                    desc.n = self._plan.grid_size(0)
            need = self._plan.workspace_bytes
            if self._ws is None or self._ws.nbytes < need:
                self._ws = DeviceArray(need, np.uint8)
        return self._plan


class PlanLauncher(object):
    """Entry of LombScargleAsyncProcess.function_tuple standing for the
    reference's ``lomb`` (NFFT) or ``lomb_dirsum`` kernel:
    ``launcher(memory, freqs)`` runs the memory's one-curve plan on its
    stream into ``memory.lsp_g``."""

    def __init__(self, name, method):
        self.name = name
        self.method = method

    def __call__(self, memory, freqs):
        plan = memory.plan_for(freqs, self.method)
        fn = "gls_lomb_execute_f64" if memory.use_double else "gls_lomb_execute"
        _lib.call(fn, plan.handle, memory.t_g.ptr, memory.yw_g.ptr, memory.w_g.ptr,
                  memory.lsp_g.ptr, memory._ws.ptr, _handle(memory.stream))

    def __repr__(self):
        return f"<libgls plan launcher {self.name}>"


# ---------------------------------------------------------------------------
# FP64 host sums (cuvarbase/lombscargle.py:337-576): the reference's CPU path,
# reached only through python_dir_sums=True.
# ---------------------------------------------------------------------------
def mhdirect_sums(t, yw, w, freq, YY, nharms=1):
    """C, S, CC, CS, SS, YC, YS of the H-harmonic model at one frequency,
    phases reduced mod 1 in FP64 (cuvarbase/lombscargle.py:337-400)."""
    phase = 2 * np.pi * ((np.asarray(t, np.float64) * freq) % 1.0)
    hs = np.arange(1, nharms + 1)
    ns = np.arange(2 * nharms + 1)
    cosn = np.cos(np.multiply.outer(ns, phase))
    sinn = np.sin(np.multiply.outer(ns, phase))
    c = cosn @ w
    s = sinn @ w
    yc = cosn[hs] @ yw
    ys = sinn[hs] @ yw
    n, k = np.meshgrid(hs, hs, indexing='ij')
    sgn = np.where(n - k == 0, 1.0, np.sign(n - k))
    cc = 0.5 * (c[n + k] + c[np.abs(n - k)])
    cs = 0.5 * (s[n + k] - sgn * s[np.abs(n - k)])
    ss = 0.5 * (c[np.abs(n - k)] - c[n + k])
    Cv, Sv = c[hs], s[hs]
    ybar = np.sum(yw)
    return (Cv, Sv, cc - np.outer(Cv, Cv), cs - np.outer(Cv, Sv), ss - np.outer(Sv, Sv),
            yc - ybar * Cv, ys - ybar * Sv)


def add_regularization(sums, amplitude_priors=None, cn0=None, sn0=None):
    """Gaussian amplitude prior on the sums (cuvarbase/lombscargle.py:403-447)."""
    C_, S, CC, CS, SS, YC, YS = sums
    D = np.zeros_like(C_)
    if amplitude_priors is not None:
        D = np.ones_like(C_) * np.power(amplitude_priors, -2)
    cn0 = np.zeros(len(C_)) if cn0 is None else cn0
    sn0 = np.zeros(len(S)) if sn0 is None else sn0
    return C_, S, CC + np.diag(D), CS, SS + np.diag(D), YC + D * cn0, YS + D * sn0


def mhgls_params_from_sums(sums, YY, ybar):
    """Amplitudes and offset (cuvarbase/lombscargle.py:450-494)."""
    C_, S, CC, CS, SS, YC, YS = sums
    H = len(C_)
    A = np.block([[CC, CS], [CS.T, SS]])
    theta = np.linalg.solve(A, np.concatenate((YC, YS)))
    cn, sn = theta[:H], theta[H:]
    return cn, sn, ybar - (np.dot(cn, C_) + np.dot(sn, S))


def mhgls_from_sums(sums, YY, ybar):
    """Periodogram power from the sums (cuvarbase/lombscargle.py:497-536)."""
    C_, S, CC, CS, SS, YC, YS = sums
    cn, sn, _ = mhgls_params_from_sums(sums, YY, ybar)
    XX = np.outer(cn, cn) * CC + 2 * np.outer(cn, sn) * CS + np.outer(sn, sn) * SS
    YX = 2 * (np.dot(cn, YC) + np.dot(sn, YS))
    return (YX - np.sum(XX)) / YY


def lomb_scargle_direct_sums(t, yw, w, freqs, YY, nharms=1, **kwargs):
    """FP64 host periodogram by direct sums (cuvarbase/lombscargle.py:539-576)."""
    sums = [add_regularization(mhdirect_sums(t, yw, w, f, YY, nharms=nharms), **kwargs)
            for f in freqs]
    ybar = np.sum(yw)
    return np.array([mhgls_from_sums(s, YY, ybar) for s in sums])




def lomb_scargle_async(memory, functions, freqs, block_size=256, use_fft=True,
                       python_dir_sums=False, transfer_to_device=True,
                       transfer_to_host=True, window=False, **kwargs):
    """GLS periodogram of one curve on ``memory.stream``
    (cuvarbase/lombscargle.py:579-689).

    ``functions`` = ((lomb, lomb_dirsum), nfft_functions) as
    LombScargleAsyncProcess builds them; the result lands in ``memory.lsp_c``
    and is valid after the stream is synchronised (``python_dir_sums``
    returns the FP64 host periodogram instead)."""
    (lomb, lomb_dirsum), _ = functions
    freqs = np.asarray(freqs, dtype=np.float64)
    assert get_k0(freqs) == memory.k0
    if transfer_to_device:
        memory.transfer_data_to_gpu()
    if python_dir_sums:
        if memory.stream is not None:
            memory.stream.synchronize()
        n = memory.n0
        t = memory.t_g.get(count=n)
        yw = memory.yw_g.get(count=n).astype(np.float64)
        w = memory.w_g.get(count=n).astype(np.float64)
        return lomb_scargle_direct_sums(t, yw, w, freqs, memory.yy)
    (lomb if use_fft else lomb_dirsum)(memory, freqs)
    if transfer_to_host:
        memory.transfer_lsp_to_cpu()
    return memory.lsp_c


class LombScargleAsyncProcess(GPUAsyncProcess):
    """GLS periodograms on the GPU (cuvarbase/lombscargle.py:692-1112).

    Keyword arguments go to NFFTAsyncProcess (``sigma`` default 4, ``m`` 8,
    ``autoset_m``, ``tol``, ``use_double``) and GPUAsyncProcess (``nstreams``,
    ``device``); ``nfft_window`` ("gaussian", the default, or "es") and
    ``window_beta`` select the spreading window; ``nharmonics`` 1-4 (the
    reference accepts only 1; above 4 raises).

    Example
    -------
    >>> proc = LombScargleAsyncProcess()
    >>> results = proc.run([(t, y, dy)])
    >>> proc.finish()
    >>> freqs, powers = results[0]
    """

    def __init__(self, *args, **kwargs):
        super(LombScargleAsyncProcess, self).__init__(*args, **kwargs)
        self.nfft_proc = NFFTAsyncProcess(*args, **kwargs)
        for name in ("_cpp_defs", "real_type", "complex_type", "block_size", "module_options",
                     "use_double"):
            setattr(self, name, getattr(self.nfft_proc, name))
        self.nfft_window = kwargs.get("nfft_window", "gaussian")
        self.window_beta = float(kwargs.get("window_beta", 0.0))
        _nfft_window(self.nfft_window)
        self.memory = None
        self.nharmonics = kwargs.get("nharmonics", 1)
        if not 1 <= self.nharmonics <= 4:
            raise Exception("nharmonics must be between 1 and 4")

    def _window_config(self):
        return dict(window=_nfft_window(self.nfft_window), window_beta=self.window_beta)

    def _compile_and_prepare_functions(self, **kwargs):
        self.nfft_proc._compile_and_prepare_functions(**kwargs)
        self._select_device()
        self.prepared_functions["lomb"] = PlanLauncher("lomb", _lib.METHOD_NFFT)
        self.prepared_functions["lomb_dirsum"] = PlanLauncher("lomb_dirsum", _lib.METHOD_DIRECT)
        self.function_tuple = (self.prepared_functions["lomb"],
                               self.prepared_functions["lomb_dirsum"])

    def _ready(self, **kwargs):
        if "lomb" not in self.prepared_functions or "lomb_dirsum" not in self.prepared_functions:
            self._compile_and_prepare_functions(**kwargs)

    # -- memory ----------------------------------------------------------------
    def memory_requirement(self, n0, nf, k0, nbatch=1, autoadjust_sigma=False, **kwargs):
        """Device bytes of ``run`` on ``nbatch`` curves of ``n0`` points and
        ``nf`` frequencies starting at k0 df: the workspace of the plan such a
        batch gets (asked of libgls; needs the device) plus the raw inputs
        (t, y, dy in FP64) and the powers.  ``autoadjust_sigma`` is accepted
        for the reference's signature and changes nothing."""
        self._select_device()
        k0 = max(int(k0), 1)
        method = _lib.METHOD_NFFT if kwargs.get("use_fft", True) else _lib.METHOD_DIRECT
        curves = [_lib.gls_curve(obs_offset=i * n0, n_obs=int(n0), nf=int(nf), f0=float(k0),
                                 df=1.0, out_offset=i * nf, ybar=0.0, yy=1.0)
                  for i in range(int(nbatch))]
        plan = Plan.lomb(curves, self.nfft_proc.sigma, self.nfft_proc.get_m(nf),
                         _lib.MODE_FLOATING_MEAN, 0.0, 0.0, _lib.INPUT_RAW, method,
                         **self._window_config())
        try:
            ws = plan.workspace_bytes
        finally:
            plan.destroy()
        out = np.dtype(self.real_type).itemsize * int(nf) * int(nbatch)
        return int(ws + 24 * int(n0) * int(nbatch) + out)

    def _memory(self, stream, m, k0, **kwargs):
        settings = dict(use_double=self.use_double, nharmonics=self.nharmonics,
                        nfft_window=self.nfft_window, window_beta=self.window_beta)
        settings.update(kwargs)
        return LombScargleMemory(self.nfft_proc.sigma, stream, m, k0=k0, **settings)

    def allocate_for_single_lc(self, t, y, dy, nf, k0=0, stream=None, **kwargs):
        mem = self._memory(stream, self.nfft_proc.get_m(nf), k0, **kwargs)
        return mem.fromdata(t=t, y=y, dy=dy, nf=nf, allocate=True, **kwargs)

    def preallocate(self, max_nobs, nlcs=1, nf=None, k0=None, freqs=None, streams=None,
                    **kwargs):
        """``nlcs`` reusable memory objects (buffered transfers of up to
        ``max_nobs`` points, ``nf`` powers) that later ``run`` calls fill in
        turn; curve i runs on ``streams[i]`` (the null stream without)."""
        if freqs is not None:
            nf, k0 = len(freqs), get_k0(freqs)
        assert nf is None or k0 is not None
        m = self.nfft_proc.get_m(nf)
        self.memory = [
            self._memory(None if streams is None else streams[i], m, k0,
                         buffered_transfer=True, n0_buffer=max_nobs, nf=nf,
                         **kwargs).allocate(**kwargs)
            for i in range(nlcs)]
        return self.memory

    def autofrequency(self, *args, **kwargs):
        return utils_autofreq(*args, **kwargs)

    def _nfreqs(self, *args, **kwargs):
        return len(self.autofrequency(*args, **kwargs))

    def allocate(self, data, nfreqs=None, k0s=None, **kwargs):
        """One LombScargleMemory per curve of ``data``, curve i on stream i;
        ``nfreqs`` / ``k0s``: a value per curve or one for all (default: the
        automatic grid's length, k0 = 1)."""
        count = len(data)
        if count:
            self._stream(count - 1)
        if nfreqs is None:
            nfreqs = [self._nfreqs(d[0], **kwargs) for d in data]
        nfreqs = np.broadcast_to(np.asarray(nfreqs, dtype=np.int64), (count,))
        k0s = np.broadcast_to(np.asarray(1 if k0s is None else k0s, dtype=np.int64), (count,))
        return [self.allocate_for_single_lc(t, y, dy, int(nf), k0=int(k0),
                                            stream=self.streams[i], **kwargs)
                for i, ((t, y, dy), nf, k0) in enumerate(zip(data, nfreqs, k0s))]

    # -- frequency grids ---------------------------------------------------------
    def _linear_grids(self, data, freqs, **kwargs):
        """One validated linear grid per curve: the automatic grid, one grid
        for all curves, or a grid per curve."""
        if freqs is None:
            grids = [self.autofrequency(d[0], **kwargs) for d in data]
        elif np.ndim(freqs[0]) == 0:
            grids = [np.asarray(freqs, dtype=np.float64)] * len(data)
        else:
            grids = [np.asarray(f, dtype=np.float64) for f in freqs]
        assert len(grids) == len(data)
        for g in grids:
            check_k0(g, k0=get_k0(g))
        return grids

    def _shared_grid(self, data, freqs, **kwargs):
        """The grid of batched_run_const_nfreq: the given one, or the
        automatic grid of the longest-baseline curve rebuilt on whole multiples
        of its spacing, df (k0 + arange(round(fmax/df) - k0))
        (cuvarbase/lombscargle.py:1041-1051)."""
        if freqs is None:
            longest = max(data, key=lambda d: np.max(d[0]) - np.min(d[0]))
            auto = self.autofrequency(longest[0], **kwargs)
            df, k0 = auto[1] - auto[0], get_k0(auto)
            freqs = df * (k0 + np.arange(int(round(np.max(auto) / df)) - k0))
        freqs = np.asarray(freqs, dtype=np.float64)
        check_k0(freqs, k0=get_k0(freqs))
        return freqs

    # -- runs ------------------------------------------------------------------
    def run(self, data, use_fft=True, memory=None, freqs=None, custom_freqs=False,
            **kwargs):
        """GLS periodograms of [(t, y, dy), ...]; returns [(freqs_i, P_i)]
        with P_i valid after finish() (cuvarbase/lombscargle.py:923-1010).

        ``custom_freqs=True`` evaluates exactly the given (arbitrarily
        spaced) frequencies by direct sums, the job of the reference's
        ``lomb_dirsum_custom_frq`` (cuvarbase/kernels/lomb.cu:167-187), which
        its Python layer never launches; the reference requires a linear grid
        (cuvarbase/lombscargle.py:939-940)."""
        self._ready(**kwargs)
        data = normalize_light_curves(data)
        if custom_freqs:
            if freqs is None:
                raise ValueError("custom_freqs=True needs explicit frequencies")
            grids = [np.asarray(freqs, np.float64)] * len(data) if np.ndim(freqs[0]) == 0 \
                else [np.asarray(f, np.float64) for f in freqs]
            if len(grids) != len(data):
                raise ValueError("need one frequency array per light curve")
            return self._run_batch(data, grids, use_fft=False, custom=True, **kwargs)
        grids = self._linear_grids(data, freqs, **kwargs)
        if self.nharmonics > 1:
            return self._run_multiharmonic(data, grids, use_fft=use_fft, memory=memory, **kwargs)
        memory = self.memory if memory is None else memory
        if memory is None:
            return self._run_batch(data, grids, use_fft=use_fft, **kwargs)
        for mem, (t, y, dy) in zip(memory, data):
            mem.set_gpu_arrays_to_zero(**kwargs)
            mem.setdata(t=t, y=y, dy=dy, **kwargs)
        opts = dict(block_size=self.block_size, use_fft=use_fft)
        opts.update(kwargs)
        funcs = (self.function_tuple, self.nfft_proc.function_tuple)
        return [(g, lomb_scargle_async(mem, funcs, g, **opts))
                for mem, g in zip(memory, grids)]

    def _run_batch(self, data, grids, use_fft=True, python_dir_sums=False,
                   transfer_to_host=True, custom=False, **kwargs):
        """All curves in one planned batch on stream 0 (FP64 host sums with
        python_dir_sums)."""
        if python_dir_sums:
            out = []
            for (t, y, dy), g in zip(data, grids):
                w = weights(dy)
                ybar = float(np.dot(w, y))
                out.append((g, lomb_scargle_direct_sums(t, w * (y - ybar), w, g,
                                                        float(np.dot(w, (y - ybar) ** 2)))))
            return out
        batch = self._launch_batch(data, grids, use_fft=use_fft, custom=custom, **kwargs)
        flat = batch.download_async(stream=self._stream(0)) if transfer_to_host else \
            pinned_zeros(max(batch.total_out, 1), batch.out_dtype)
        return list(zip(grids, batch.split(flat)))

    def _run_multiharmonic(self, data, grids, use_fft=True, memory=None, **kwargs):
        """nharmonics > 1: harmonic NFFTs of one plan and an FP64 solve per
        frequency (engine.MultiHarmonicBatch)."""
        if not use_fft or memory is not None or kwargs.get("python_dir_sums", False):
            raise NotImplementedError("nharmonics > 1 runs on the batched NFFT path only "
                                      "(use_fft=True, no preallocated memory)")
        if not kwargs.get("floating_mean", True) or kwargs.get("window", False):
            raise NotImplementedError("nharmonics > 1 is the floating-mean model only")
        from .engine import MultiHarmonicBatch

        m = max((self.nfft_proc.get_m(len(g)) for g in grids), default=self.nfft_proc.m)
        batch = MultiHarmonicBatch(data, grids, nharmonics=self.nharmonics,
                                   sigma=self.nfft_proc.sigma, m=m,
                                   amplitude_prior=kwargs.get("amplitude_prior", None),
                                   use_double=self.use_double)
        return list(zip(grids, batch.run(stream=self._stream(0))))

    def _launch_batch(self, data, grids, use_fft=True, custom=False, **kwargs):
        """Plan, upload and execute one LombBatch on stream 0 (no download);
        its buffers stay alive until finish()."""
        m = max((self.nfft_proc.get_m(len(g)) for g in grids), default=self.nfft_proc.m)
        batch = LombBatch(data, grids, sigma=self.nfft_proc.sigma, m=m,
                          mode=_mode(kwargs.get("floating_mean", True),
                                     kwargs.get("window", False)),
                          amplitude_prior=kwargs.get("amplitude_prior", None),
                          use_fft=use_fft, use_double=self.use_double, input_kind="raw",
                          custom_freqs=custom, **self._window_config())
        stream = self._stream(0)
        batch.upload(stream=stream)
        batch.execute(stream=stream)
        self._keep_alive.append(batch)
        return batch

    def _significance(self, curve, z, fmax):
        """1 - Baluev FAP of peak power z of a centred curve."""
        if not np.isfinite(z):
            return np.nan
        return 1. - fap_baluev(curve[0], curve[2], np.array([z]), fmax)[0]

    def batched_run_const_nfreq(self, data, batch_size=10, use_fft=True, freqs=None,
                                only_return_best_freqs=False, ignore_freq_mask=None,
                                **kwargs):
        """Periodograms of many curves on one shared grid
        (cuvarbase/lombscargle.py:1012-1112): [(freqs, P_i)], or with
        ``only_return_best_freqs`` the best frequency of each curve (ignoring
        the frequencies where ``ignore_freq_mask`` is True) and its
        significance 1 - fap_baluev at that power and the highest unmasked
        frequency.  Slices of ``batch_size`` curves run as one batch each;
        best frequencies are selected on the device (one value per curve comes
        back) unless the run needs host powers (use_double, python_dir_sums,
        preallocated memory)."""
        self._ready(**kwargs)
        data = normalize_light_curves(data)
        grid = self._shared_grid(data, freqs, **kwargs)
        keep = np.ones(len(grid), dtype=bool) if ignore_freq_mask is None else \
            ~np.asarray(ignore_freq_mask, dtype=bool)
        fmax = float(np.max(grid[keep]))
        on_device = only_return_best_freqs and not self.use_double and \
            not kwargs.get("python_dir_sums", False) and self.memory is None
        best, significance, powers = [], [], []
        for lo in range(0, len(data), batch_size):
            part = data[lo:lo + batch_size]
            if on_device:
                batch = self._launch_batch(part, [grid] * len(part), use_fft=use_fft, **kwargs)
                skip = None if ignore_freq_mask is None else \
                    np.tile(~keep, len(part)).astype(np.uint8)
                top_p, top_i = batch.topk(1, mask=skip, stream=self._stream(0))
                for curve, z, k in zip(part, top_p[:, 0], top_i[:, 0]):
                    best.append(grid[k] if k >= 0 else np.nan)
                    significance.append(self._significance(curve, float(z) if k >= 0 else np.nan,
                                                           fmax))
            else:
                for curve, (_, p) in zip(part, self.run(part, freqs=grid, use_fft=use_fft,
                                                        **kwargs)):
                    self.finish()
                    if only_return_best_freqs:
                        k = int(np.argmax(np.where(keep, p, -np.inf)))
                        best.append(grid[k])
                        significance.append(self._significance(curve, float(p[k]), fmax))
                    else:
                        powers.append(np.array(p))
            self.finish()
        if only_return_best_freqs:
            return best, significance
        return [(grid, p) for p in powers]


def fap_baluev(t, dy, z, fmax, d_K=3, d_H=1, use_gamma=True):
    """Baluev (2008) false-alarm probability of periodogram values ``z``
    (cuvarbase/lombscargle.py:1115-1184), including the reference's form of
    the first exponent factor, (z/pi)^0.5 * (d-1) (:1177)."""
    t = np.asarray(t, dtype=np.float64)
    N = len(t)
    d = d_K - d_H
    N_K = N - d_K
    N_H = N - d_H
    g = (0.5 * N_H) ** (0.5 * (d - 1))
    if use_gamma:
        g = np.exp(gammaln(0.5 * N_H) - gammaln(0.5 * (N_K + 1)))
    w = np.power(np.asarray(dy, dtype=np.float64), -2)
    tbar = np.dot(w, t) / np.sum(w)
    Dt = np.dot(w, np.power(t - tbar, 2)) / np.sum(w)
    Teff = np.sqrt(4 * np.pi * Dt)
    A = (2 * np.pi ** 1.5) * fmax * Teff
    z = np.asarray(z, dtype=np.float64)
    eZ1 = (z / np.pi) ** 0.5 * (d - 1)
    eZ2 = (1 - z) ** (0.5 * (N_K - 1))
    tau = (g * A / (2 * np.pi)) * eZ1 * eZ2
    Psing = 1 - (1 - z) ** (0.5 * N_K)
    return 1 - Psing * np.exp(-tau)


def lomb_scargle_simple(t, y, dy, **kwargs):
    """One-curve convenience call (cuvarbase/lombscargle.py:1187-1204), passing
    ``dy`` where the reference passed normalised weights."""
    proc = LombScargleAsyncProcess()
    results = proc.run([(t, y, dy)], **kwargs)
    proc.finish()
    freqs, powers = results[0]
    return freqs, np.array(powers)
