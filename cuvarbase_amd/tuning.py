"""Choice of the NFFT oversampling sigma and Gaussian half-width m.

The band-centred adjoint NFFT with Gaussian width b = 2 sigma m/((2 sigma - 1) pi)
(cuvarbase/cunfft.py:56-57) has two error terms in the six GLS sums when the
grid and its FFT are held in FP32:

* the window's alias / truncation error, which falls with m:
  ~ C_ALIAS * exp(-A m),  A = pi (2 sigma - 2) / (2 sigma - 1);
* FP32 round-off of the spread grid and its FFT, which the deconvolution
  exp(b pi^2 (l/n)^2) multiplies by up to exp(b pi^2 / (4 sigma^2)) at the band
  edges (|l/n| = 1/(2 sigma)), i.e. ~ C_FP32 * exp(B m),
  B = pi / (2 sigma (2 sigma - 1)).

A larger m therefore stops helping once the second term dominates, and a small
sigma (sharp band edges) makes the second term explode: sigma = 1.2, m = 18
would be accurate to 1e-7 in exact arithmetic but loses every digit in FP32
(edge factor ~2e7).  The constants were fitted to the power errors |dP| of an
FP32 emulation of the kernels' arithmetic (FP32 spreading, complex64 FFT,
FP32 deconvolution, FP64 assembly) against an FP64 sigma=4, m=14 NFFT on
survey curves: for example (sigma, m) = (2, 8) -> 2.5e-7, (1.5, 8) -> 3.3e-6,
(1.4, 8) -> 7.0e-6, (1.3, 7) -> 1.5e-4, (1.25, 6) -> 5.3e-4, (1.2, 18) -> 1.0.

The grid work is proportional to sigma and the spreading work to N (2m + 1),
so ``choose_nfft_params`` returns the smallest sigma (then the smallest m)
whose predicted |dP| is below tol / SAFETY.  ``calibrate`` checks a choice on
the GPU against the direct-sum path.
"""
from __future__ import annotations

import math

import numpy as np

SIGMAS = (1.25, 1.3, 1.35, 1.4, 1.45, 1.5, 1.6, 1.75, 2.0, 2.5, 3.0, 4.0)
M_RANGE = range(4, 17)
SAFETY = 10.0        # required margin of the predicted |dP| below the target
C_ALIAS = 0.3
C_FP32 = 5e-8


def alias_error(sigma, m):
    return C_ALIAS * math.exp(-math.pi * m * (2 * sigma - 2) / (2 * sigma - 1))


def fp32_error(sigma, m):
    return C_FP32 * math.exp(math.pi * m / (2 * sigma * (2 * sigma - 1)))


def predicted_power_error(sigma, m):
    """Model max |dP| of the FP32 NFFT path with (sigma, m)."""
    return alias_error(sigma, m) + fp32_error(sigma, m)


def deconvolution_edge_factor(sigma, m):
    """Largest deconvolution factor exp(b pi^2 / (4 sigma^2))."""
    return math.exp(math.pi * m / (2 * sigma * (2 * sigma - 1)))


def candidates(tol=1e-4, m_max=16):
    """Admissible (sigma, m) pairs, cheapest first."""
    out = []
    for sigma in SIGMAS:
        for m in M_RANGE:
            if m <= m_max and predicted_power_error(sigma, m) * SAFETY <= tol:
                out.append((sigma, m))
                break
    return out


def choose_nfft_params(tol=1e-4, m_max=16):
    """Cheapest (sigma, m) with predicted |dP| <= tol / SAFETY."""
    c = candidates(tol, m_max)
    return c[0] if c else (2.0, 8)


# Exponential-of-semicircle window psi(z) = exp(beta (sqrt(1 - z^2) - 1)) on
# 2m + 1 cells, beta = gamma pi (2m + 1)(1 - 1/(2 sigma)).  Its transform is
# flat across the band, so the deconvolution amplifies FP32 round-off far
# less than the Gaussian's exp(b pi^2 xi^2): the same accuracy needs a smaller
# sigma and m.  Entries (sigma, m): (gamma, max |dP|) measured with the FP32
# emulation described above on survey curves, including the hardest ones
# (50-100 points, whose lowest frequencies are ill-conditioned and are taken
# by the FP64 direct-sum fix-up of the library: f (t_max - t_min) < 3).
ES_MEASURED = {
    (1.25, 6): (1.05, 3.6e-5),
    (1.3, 5): (1.00, 2.3e-5),
    (1.3, 6): (1.00, 1.06e-5),
    (1.35, 5): (1.05, 8.9e-6),
}
# relative cost per evaluation: grid cells (sigma) plus spreading taps
SPREAD_COST_PER_TAP = 0.015


def es_beta(sigma, m, gamma=1.0):
    return gamma * math.pi * (2 * m + 1) * (1.0 - 1.0 / (2.0 * sigma))


def config_cost(sigma, m):
    return sigma * (1.0 + SPREAD_COST_PER_TAP * (2 * m + 1))


def choose_nfft_config(tol=1e-4):
    """Cheapest NFFT configuration for |dP| <= tol / SAFETY among the Gaussian
    model candidates and the measured ES entries: a dict with sigma, m,
    window (0 Gaussian, 1 ES) and window_beta."""
    best = None
    for sigma, m in candidates(tol):
        c = config_cost(sigma, m)
        if best is None or c < best[0]:
            best = (c, dict(sigma=sigma, m=m, window=0, window_beta=0.0))
    for (sigma, m), (gamma, err) in ES_MEASURED.items():
        if err * SAFETY <= tol:
            c = config_cost(sigma, m)
            if best is None or c < best[0]:
                best = (c, dict(sigma=sigma, m=m, window=1,
                                window_beta=es_beta(sigma, m, gamma)))
    return best[1] if best else dict(sigma=2.0, m=8, window=0, window_beta=0.0)


def calibrate(data, freqs, tol=1e-4, probes=4, nprobe_freqs=256, seed=0):
    """Measured choice on the GPU: returns (sigma, m, measured max |dP|).

    For each admissible candidate (cheapest first) the NFFT path is compared
    with the direct-sum path (FP64 phases, ~1e-7 from FP64) on ``probes``
    curves at ``nprobe_freqs`` random frequencies each; the first whose error
    is below tol / 10 is returned."""
    from .engine import LombBatch

    rng = np.random.default_rng(seed)
    idx = rng.choice(len(data), size=min(probes, len(data)), replace=False)
    sub = [data[i] for i in idx]
    sub_f = [freqs[i] for i in idx]
    sel = [rng.choice(len(f), size=min(nprobe_freqs, len(f)), replace=False) for f in sub_f]
    ref = LombBatch(sub, sub_f, use_fft=False).run()
    for sigma, m in candidates(tol * SAFETY / 10.0):
        got = LombBatch(sub, sub_f, sigma=sigma, m=m).run()
        err = max(float(np.max(np.abs(g[s].astype(np.float64) - r[s])))
                  for g, r, s in zip(got, ref, sel))
        if err * 10.0 <= tol:
            return sigma, m, err
    return 2.0, 8, float("nan")
