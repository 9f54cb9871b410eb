"""FP64 adjoint NFFT by Gaussian gridding (test infrastructure only).

Restates the NFFT arithmetic of cuvarbase/kernels/cunfft.cu:36-248 and the launch
sequence of nfft_adjoint_async (cuvarbase/cunfft.py:163-308) in FP64 NumPy.

Two band layouts are provided:

* ``nfft_adjoint`` -- the layout the HIP library uses: the band of K output
  frequencies f_k = f0 + k df is centred by modulating each sample with
  exp(2 pi i Fc (t - tref)), Fc = f0 + (K//2) df, so the outputs occupy the
  FFT indices -K//2 .. K - K//2 - 1 (mod n).  Alias error ~ exp(-pi^2 b (1 - 1/s))
  with s = n/K.
* ``nfft_adjoint_reference_style`` -- the reference's own one-sided band: real
  gridding (fast_gaussian_grid, cunfft.cu:115-163), grid shift by kappa = f0/df
  (nfft_shift, cunfft.cu:59-82), unnormalised inverse FFT (cunfft.py:288), then
  phase + Gaussian deconvolution (normalize, cunfft.cu:214-248).  It is kept to
  document the accuracy of the reference layout; nothing is graded against it.

Pins (tests/test_oracle_nfft.py):

==================================  ==========================================
function                            pinned by
==================================  ==========================================
gaussian_grid_reference             brute-force cell-by-cell gridding loop
                                    (``gaussian_grid_bruteforce``) on tiny
                                    inputs; the scalar recurrence of
                                    cuvarbase/tests/test_nfft.py:57-87
                                    (q1*q2^k*q3[k] == psi)
FFT convention                      n * numpy.fft.ifft == explicit DFT with
                                    + sign (cuvarbase/tests/test_nfft.py:186-200)
nfft_adjoint                        direct DFT sum(v exp(2 pi i f t)) in FP64
                                    with mod-1 phase reduction, error below
                                    the Gaussian alias/truncation bound
nfft_adjoint_reference_style        direct DFT at sigma=5 (the reference's test
                                    setting, cuvarbase/tests/test_nfft.py:24)
gls_nfft                            oracle.gls.gls_direct (|dP| < 1e-6)
==================================  ==========================================
"""
from __future__ import annotations

import numpy as np

from . import gls as _gls


def gauss_b(sigma, m):
    """Gaussian width b = 2 sigma m / ((2 sigma - 1) pi)  (cuvarbase/cunfft.py:56-57)."""
    return 2.0 * sigma * m / ((2.0 * sigma - 1.0) * np.pi)


def _frac(x):
    return x - np.floor(x)


def _psi(d, b):
    """Gaussian window exp(-d^2/b)/sqrt(pi b)  (cuvarbase/kernels/cunfft.cu:36-38)."""
    return np.exp(-(d * d) / b) / np.sqrt(np.pi * b)


def spread(s, v, n, m, b):
    """Spread complex samples v at cell coordinates n*s (s in [0,1)) onto a
    periodic grid of n cells with 2m+1 Gaussian taps per sample.

    Tap cells U+k, U = floor(n s) - m, k = 0..2m, weight psi(n s - U - k):
    the tap set of fast_gaussian_grid (cuvarbase/kernels/cunfft.cu:146-161).
    """
    y = n * np.asarray(s, dtype=np.float64)
    fl = np.floor(y)
    U = fl.astype(np.int64) - m
    fr = y - fl
    grid = np.zeros(n, dtype=np.complex128)
    for k in range(2 * m + 1):
        wgt = _psi(m + fr - k, b)
        np.add.at(grid, (U + k) % n, v * wgt)
    return grid


def nfft_adjoint(t, v, f0, df, K, n=None, sigma=2.0, m=8, b=None, tref=None,
                 centred=True, apply_phase=True):
    """V_k ~= sum_j v_j exp(2 pi i (f0 + k df) t_j), k = 0..K-1, in FP64.

    With apply_phase=False the result is sum_j v_j exp(2 pi i f_k (t_j - tref))
    (the time origin moved to tref; GLS power does not depend on it).
    """
    t = np.asarray(t, dtype=np.float64)
    v = np.asarray(v, dtype=np.complex128)
    tref = float(np.min(t)) if tref is None else float(tref)
    if n is None:
        n = int(np.ceil(sigma * K))
    if b is None:
        b = gauss_b(n / K, m)
    kc = K // 2 if centred else 0
    Fc = f0 + kc * df
    dt = t - tref
    s = _frac(dt * df)
    vv = v * np.exp(2j * np.pi * _frac(Fc * dt))
    g = spread(s, vv, n, m, b)
    G = np.fft.ifft(g) * n
    lp = np.arange(K) - kc
    V = G[lp % n] * np.exp(b * (np.pi * lp / n) ** 2)
    if apply_phase:
        fk = f0 + np.arange(K) * df
        V = V * np.exp(2j * np.pi * _frac(fk * tref))
    return V


def direct_dft(t, v, freqs):
    """sum_j v_j exp(2 pi i f t_j) in FP64 with mod-1 phase reduction
    (cuvarbase/tests/test_nfft.py:31-35)."""
    t = np.asarray(t, dtype=np.float64)
    v = np.asarray(v, dtype=np.complex128)
    out = np.empty(len(freqs), dtype=np.complex128)
    for a in range(0, len(freqs), 256):
        ph = _frac(np.multiply.outer(np.asarray(freqs[a:a + 256]), t))
        out[a:a + 256] = np.exp(2j * np.pi * ph) @ v
    return out


# ----------------------------------------------------------------------------
# the reference's one-sided layout
# ----------------------------------------------------------------------------
def gaussian_grid_reference(t, y, nf, sigma, m, samples_per_peak=1.0):
    """Real Gaussian grid of size int(sigma*nf) as fast_gaussian_grid makes it
    (cuvarbase/kernels/cunfft.cu:84-163, sizes cuvarbase/cunfft.py:87):
    x = (t - tmin)/(spp (tmax - tmin)), grid[(U+k) mod ng] += y psi(ng x - U - k)."""
    t = np.asarray(t, dtype=np.float64)
    ng = int(sigma * nf)
    b = gauss_b(sigma, m)
    x = (t - t.min()) / (samples_per_peak * (t.max() - t.min()))
    return spread(_frac(x), np.asarray(y, dtype=np.complex128), ng, m, b).real


def gaussian_grid_bruteforce(t, y, nf, sigma, m, samples_per_peak=1.0):
    """Cell-by-cell restatement: grid[u] = sum_j y_j psi(d) over every d =
    ng x_j - u + p ng (p integer) that lies in [fr_j - m, fr_j + m]."""
    t = np.asarray(t, dtype=np.float64)
    ng = int(sigma * nf)
    b = gauss_b(sigma, m)
    x = (t - t.min()) / (samples_per_peak * (t.max() - t.min()))
    grid = np.zeros(ng)
    for u in range(ng):
        acc = 0.0
        for xj, yj in zip(x, y):
            yy = ng * (xj - np.floor(xj))
            fr = yy - np.floor(yy)
            d0 = yy - u
            for p in range(-2, 3):
                d = d0 + p * ng
                if fr - m - 1e-9 <= d <= fr + m + 1e-9:
                    acc += yj * _psi(d, b)
        grid[u] = acc
    return grid


def nfft_adjoint_reference_style(t, y, nf, minimum_frequency=0.0,
                                 samples_per_peak=1.0, sigma=4, m=8):
    """FP64 restatement of nfft_adjoint_async's sequence (one-sided band).

    precompute_psi + fast_gaussian_grid -> nfft_shift (only if |f0| > 1e-9,
    cuvarbase/cunfft.py:276) -> n*ifft -> normalize.  Returns ghat[k] ~
    sum_j y_j exp(2 pi i (f0 + k df) t_j), df = 1/(spp (tmax - tmin)).
    """
    t = np.asarray(t, dtype=np.float64)
    tmin, tmax = t.min(), t.max()
    ng = int(sigma * nf)
    b = gauss_b(sigma, m)
    g = gaussian_grid_reference(t, y, nf, sigma, m, samples_per_peak).astype(
        np.complex128)
    sT = samples_per_peak * (tmax - tmin)
    kappa = minimum_frequency * sT
    if abs(minimum_frequency) > 1e-9:
        g = g * np.exp(2j * np.pi * np.arange(ng) * kappa / ng)
    G = np.fft.ifft(g) * ng
    k = np.arange(nf)
    n0 = (tmin / sT) * ng
    G = G[:nf] * np.exp(2j * np.pi * n0 * (kappa + k) / ng)
    return G * np.exp(b * (np.pi * (kappa + k) / ng) ** 2)


# ----------------------------------------------------------------------------
# NFFT-based GLS (the CPU baseline path)
# ----------------------------------------------------------------------------
def gls_nfft(t, y, dy, freqs, sigma=2.0, m=8, mode=_gls.FLOATING_MEAN,
             amplitude_prior=None, normalize=True, n=None, b=None):
    """GLS power at the linear grid ``freqs`` from three FP64 adjoint NFFTs:
    SYW at f, SW at f and SW at 2f (the sums of cuvarbase/kernels/lomb.cu:204-216).
    """
    if normalize:
        t, y, dy = _gls.normalize_light_curve(t, y, dy)
    w, yw, ybar, YY = _gls.prepare(y, dy)
    freqs = np.asarray(freqs, dtype=np.float64)
    f0, df, K = float(freqs[0]), float(freqs[1] - freqs[0]), len(freqs)
    tref = float(np.min(t))
    kw = dict(n=n, sigma=sigma, m=m, b=b, tref=tref, apply_phase=False)
    syw = nfft_adjoint(t, yw, f0, df, K, **kw)
    w1 = nfft_adjoint(t, w, f0, df, K, **kw)
    w2 = nfft_adjoint(t, w, 2 * f0, 2 * df, K, **kw)
    reg = None
    if amplitude_prior is not None:
        lam = float(amplitude_prior) ** -2
        reg = (lam, lam)
    return _gls.power_from_sums(w1.real, w1.imag, w2.real, w2.imag,
                                syw.real, syw.imag, YY, ybar, mode=mode,
                                reg=reg)
