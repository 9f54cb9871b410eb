"""Seeded synthetic light-curve workloads used by the benchmark and tests.

Each generator returns plain NumPy data: ``[(t, y, dy), ...]`` plus one
frequency grid per curve.  Seeds are fixed per workload so every run sees the
same inputs.

=========  ===================================================================
workload   description
=========  ===================================================================
single     N=1000 sorted times on [0, 10] d, y = sin(2 pi t/0.73) + N(0, 0.5^2),
           dy = 0.5, 10,000 linear frequencies 0.05..50 /d
survey     5,000 curves, N log-uniform in [50, 5000], baselines U[30, 3650] d
           with seasonal gaps (0.6 of each 365-d cycle observable) and nightly
           jitter; y = A sin(2 pi t/P + phi) + N(0, s^2); df = 1/(5 T),
           fmax = 24 /d
long       one curve, N = 200,000 over 3650 d, P = 0.0371 d, noise 0.1,
           df = 1/(5 T), fmax = 100 /d (~1.8M frequencies)
stream     request generator: N log-uniform in [100, 50,000], T U[30, 3650] d,
           grid as survey; Poisson arrival times at a given rate
short      100,000 curves, N uniform in [20, 300] over 90 d, 20,000
           frequencies 0.01..40 /d; half pure noise, half sinusoid + noise
=========  ===================================================================
"""
from __future__ import annotations

import numpy as np


def _seasonal_times(rng, n, T):
    """n times over a baseline T with 0.6 of every 365-day cycle observable
    and +-0.15 d nightly jitter, sorted."""
    nights = np.arange(int(np.ceil(T)))
    ok = nights[(nights % 365) < 0.6 * 365]
    d = rng.choice(ok, size=n, replace=True).astype(np.float64)
    t = d + rng.uniform(-0.15, 0.15, size=n)
    return np.sort(t)


def survey_grid(t, fmax=24.0, spp=5.0):
    """df = 1/(spp * baseline); f = df * (1 + arange(int(fmax/df) - 1))."""
    df = 1.0 / (spp * (t.max() - t.min()))
    nf = int(fmax / df) - 1
    return df * (1.0 + np.arange(nf))


def single(seed=1):
    rng = np.random.default_rng(seed)
    n = 1000
    t = np.sort(rng.uniform(0.0, 10.0, n))
    y = np.sin(2 * np.pi * t / 0.73) + rng.normal(0.0, 0.5, n)
    dy = np.full(n, 0.5)
    freqs = np.linspace(0.05, 50.0, 10000)
    return [(t, y, dy)], [freqs]


def survey_curve(rng, n_range=(50, 5000), T_range=(30.0, 3650.0), fmax=24.0):
    n = int(np.floor(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
    T = rng.uniform(*T_range)
    t = _seasonal_times(rng, n, T)
    P = np.exp(rng.uniform(np.log(0.05), np.log(100.0)))
    A = rng.uniform(0.05, 1.0)
    s = rng.uniform(0.01, 0.5)
    phi = rng.uniform(0.0, 2 * np.pi)
    y = A * np.sin(2 * np.pi * t / P + phi) + rng.normal(0.0, s, n)
    dy = np.full(n, s)
    return (t, y, dy), survey_grid(t, fmax=fmax), P


def survey(ncurves=5000, seed=2):
    rng = np.random.default_rng(seed)
    data, freqs, periods = [], [], []
    for _ in range(ncurves):
        d, f, P = survey_curve(rng)
        data.append(d)
        freqs.append(f)
        periods.append(P)
    return data, freqs


def long(seed=3, n=200_000, T=3650.0, fmax=100.0):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, T, n))
    y = np.sin(2 * np.pi * t / 0.0371 + rng.uniform(0, 2 * np.pi)) + rng.normal(0.0, 0.1, n)
    dy = np.full(n, 0.1)
    return [(t, y, dy)], [survey_grid(t, fmax=fmax)]


def stream_requests(rate, duration=None, nreq=None, seed=4):
    """Poisson arrivals: returns (arrival_times, data, freqs)."""
    rng = np.random.default_rng(seed)
    if nreq is None:
        nreq = int(rate * duration)
    gaps = rng.exponential(1.0 / rate, size=nreq)
    arrivals = np.cumsum(gaps)
    data, freqs = [], []
    for _ in range(nreq):
        d, f, _P = survey_curve(rng, n_range=(100, 50_000))
        data.append(d)
        freqs.append(f)
    return arrivals, data, freqs


def short(ncurves=100_000, seed=5, nf=20000):
    rng = np.random.default_rng(seed)
    freqs = np.linspace(0.01, 40.0, nf)
    data = []
    for i in range(ncurves):
        n = int(rng.integers(20, 301))
        t = np.sort(rng.uniform(0.0, 90.0, n))
        y = rng.normal(0.0, 1.0, n)
        if i % 2 == 1:
            A = rng.uniform(0.5, 2.0)
            P = np.exp(rng.uniform(np.log(0.05), np.log(45.0)))
            y = y + A * np.sin(2 * np.pi * t / P + rng.uniform(0, 2 * np.pi))
        data.append((t, y, np.ones(n)))
    return data, [freqs] * ncurves
