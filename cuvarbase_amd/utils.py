"""Host utilities of the GLS path: observation weights, the automatic
frequency grid and the centring of light curves.

These keep the names, arguments and results of cuvarbase/utils.py:10-13
(``weights``), :46-97 (``autofrequency``), :106-110 (``get_autofreqs``) and
:113-146 (``normalize_light_curves``).
"""
from __future__ import annotations

import numpy as np

# the keyword arguments that shape the automatic grid
GRID_KEYWORDS = ("minimum_frequency", "maximum_frequency", "nyquist_factor",
                 "samples_per_peak")


def weights(err):
    """Inverse-variance weights with unit sum: err_i^-2 / sum_j err_j^-2."""
    inv_var = np.reciprocal(np.square(np.asarray(err, dtype=np.float64)))
    return inv_var / np.sum(inv_var)


def autofrequency(t, nyquist_factor=5, samples_per_peak=5,
                  minimum_frequency=None, maximum_frequency=None, **kwargs):
    """Linear grid on multiples of df = 1 / (baseline * samples_per_peak).

    The grid starts at the multiple ``max(1, int(minimum_frequency / df))``
    (1 without a minimum) and holds ``int(maximum_frequency / df)`` minus that
    many frequencies, or ``int(samples_per_peak * nyquist_factor * N / 2)``
    without a maximum -- the "average Nyquist" heuristic of N points.
    Extra keyword arguments are ignored."""
    t = np.asarray(t)
    df = 1.0 / ((np.max(t) - np.min(t)) * samples_per_peak)
    first = 1 if minimum_frequency is None else max(1, int(minimum_frequency / df))
    if maximum_frequency is None:
        count = int(0.5 * samples_per_peak * nyquist_factor * len(t))
    else:
        count = int(maximum_frequency / df) - first
    return df * np.arange(first, first + count)


def get_autofreqs(t, **kwargs):
    """``autofrequency(t)`` driven by the grid keywords found in ``kwargs``."""
    return autofrequency(t, **{k: kwargs[k] for k in GRID_KEYWORDS if k in kwargs})


def normalize_light_curves(data):
    """New tuples with t and y centred on their NaN-ignoring means; any further
    columns (dy, ...) are copied unchanged.  The caller's arrays are never
    modified."""
    centred = []
    for columns in data:
        cols = [np.array(c, copy=True) for c in columns]
        for j in range(min(2, len(cols))):
            cols[j] = cols[j] - np.nanmean(cols[j])
        centred.append(tuple(cols))
    return centred
