"""Choice of the NFFT oversampling sigma and Gaussian half-width m.

The band-centred adjoint NFFT with Gaussian width b = 2 sigma m/((2 sigma - 1) pi)
has relative error ~ eps(sigma, m) = exp(-pi m (1 - 1/sigma) / (1 - 1/(2 sigma)))
in the six sums (alias term and deconvolution-amplified truncation term
balanced).  Power errors follow as |dP| ~ C eps with C ~ 10 (measured on
survey curves against FP64 sums).  The grid work is proportional to sigma, the
spreading work to N (2m + 1), which is negligible next to the grid for the
workloads here, so the cheapest admissible choice is the smallest sigma whose m
(<= 24) meets the target.

``choose_nfft_params`` applies that model.  ``calibrate`` checks a choice
empirically on the GPU: the NFFT path against the GPU direct-sum path (FP64
phases, FP32 rotation, ~1e-7 from FP64) on probe curves and frequencies, and
walks to the next candidate if the measured error exceeds the target.
"""
from __future__ import annotations

import math

import numpy as np

# (sigma, m) candidates ordered by grid cost
CANDIDATES = [(1.2, 18), (1.25, 16), (1.35, 14), (1.5, 12), (2.0, 8), (2.0, 12), (4.0, 12)]
SAFETY = 100.0       # required margin of the predicted |dP| below the target
C_POWER = 10.0       # |dP| ~ C_POWER * eps


def sum_error(sigma, m):
    """Model relative error of the NFFT sums."""
    return math.exp(-math.pi * m * (1.0 - 1.0 / sigma) / (1.0 - 1.0 / (2.0 * sigma)))


def predicted_power_error(sigma, m):
    return C_POWER * sum_error(sigma, m)


def choose_nfft_params(tol=1e-4, m_max=24):
    """Cheapest (sigma, m) with predicted |dP| <= tol / SAFETY."""
    for sigma, m in CANDIDATES:
        if m <= m_max and predicted_power_error(sigma, m) * SAFETY <= tol:
            return sigma, m
    return CANDIDATES[-1]


def calibrate(data, freqs, tol=1e-4, probes=4, nprobe_freqs=256, seed=0):
    """Measured choice on the GPU: returns (sigma, m, measured max |dP|).

    For each candidate (cheapest first) that the model admits, the NFFT path is
    compared with the direct-sum path on ``probes`` curves at
    ``nprobe_freqs`` random frequencies each."""
    from .engine import LombBatch

    rng = np.random.default_rng(seed)
    idx = rng.choice(len(data), size=min(probes, len(data)), replace=False)
    sub = [data[i] for i in idx]
    sub_f = [freqs[i] for i in idx]
    sel = [rng.choice(len(f), size=min(nprobe_freqs, len(f)), replace=False) for f in sub_f]
    ref = LombBatch(sub, sub_f, use_fft=False).run()
    for sigma, m in CANDIDATES:
        if predicted_power_error(sigma, m) > tol:
            continue
        got = LombBatch(sub, sub_f, sigma=sigma, m=m).run()
        err = max(float(np.max(np.abs(g[s].astype(np.float64) - r[s])))
                  for g, r, s in zip(got, ref, sel))
        if err * 10.0 <= tol:
            return sigma, m, err
    return 4.0, 12, float("nan")
