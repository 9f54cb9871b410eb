"""Host utilities of the GLS path (cuvarbase/utils.py:10-146)."""
from __future__ import annotations

from copy import deepcopy

import numpy as np


def weights(err):
    """Observation weights from uncertainties, normalised to sum 1
    (cuvarbase/utils.py:10-13)."""
    w = np.power(err, -2)
    return w / sum(w)


def tophat_window(t, t0, d):
    """cuvarbase/utils.py:35-38."""
    w_window = np.zeros_like(t)
    w_window[np.absolute(t - t0) < d] += 1.
    return w_window / max(w_window)


def gaussian_window(t, t0, d):
    """cuvarbase/utils.py:41-43."""
    w_window = np.exp(-0.5 * np.power(t - t0, 2) / (d * d))
    return w_window / (1. if len(w_window) == 0 else max(w_window))


def autofrequency(t, nyquist_factor=5, samples_per_peak=5,
                  minimum_frequency=None, maximum_frequency=None, **kwargs):
    """Heuristic linear frequency grid ``df * (nf0 + arange(Nf))``
    (cuvarbase/utils.py:46-97): df = 1/(baseline * samples_per_peak);
    nf0 = max(1, int(minimum_frequency/df)); Nf = int(maximum_frequency/df) - nf0
    or int(0.5 * samples_per_peak * nyquist_factor * len(t))."""
    baseline = max(t) - min(t)
    n_samples = len(t)
    df = 1. / (baseline * samples_per_peak)
    nf0 = 1
    if minimum_frequency is not None:
        nf0 = max([nf0, int(minimum_frequency / df)])
    if maximum_frequency is not None:
        Nf = int(maximum_frequency / df) - nf0
    else:
        Nf = int(0.5 * samples_per_peak * nyquist_factor * n_samples)
    return df * (nf0 + np.arange(Nf))


def dphase(dt, freq):
    """Phase distance in [0, 1/2] (cuvarbase/utils.py:100-103)."""
    dph = dt * freq - np.floor(dt * freq)
    return dph if dph < 0.5 else 1 - dph


def get_autofreqs(t, **kwargs):
    """autofrequency with only its own keyword arguments (cuvarbase/utils.py:106-110)."""
    keys = ('minimum_frequency', 'maximum_frequency', 'nyquist_factor', 'samples_per_peak')
    return autofrequency(t, **{k: v for k, v in kwargs.items() if k in keys})


def normalize_light_curves(data):
    """Copy each (t, y, ...) tuple and subtract the nanmean of t and y
    (cuvarbase/utils.py:113-146).  The caller's arrays are not modified."""
    data = deepcopy(data)
    for i, lc in enumerate(data):
        out = []
        for j in range(len(lc)):
            if j < 2:
                out.append((np.asarray(lc[j]) - np.nanmean(lc[j])).copy())
            else:
                out.append(np.asarray(lc[j]).copy())
        data[i] = tuple(out)
    return data
