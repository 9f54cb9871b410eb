"""Pins of oracle.gls.fap_baluev (cuvarbase/lombscargle.py:1115-1184) by
mathematics and simulation, not by a second transcription of its formula.

fap_baluev(z) = 1 - Psing(z) exp(-tau(z)), Psing the single-frequency CDF of
the GLS power and tau Baluev's (2008) Davies bound for the number of upcrossings
over [0, fmax].  Each check below fails for a plausible transcription mistake:

* single-frequency tail: for Gaussian noise and a fixed frequency, the
  floating-mean GLS power of N points is Beta(1, (N - 3)/2), so
  P(z > z0) = (1 - z0)^((N-3)/2).  fap_baluev with fmax -> 0 (tau -> 0) must
  reproduce the simulated tail (an exponent built from N - 1 or N would not:
  0.14 vs 0.20 at N = 12, z0 = 0.3);
* the whole FAP against the simulated distribution of the maximum power over a
  dense grid to fmax, pure noise, uniform random sampling (whose weighted time
  variance makes Teff = sqrt(4 pi Dt) ~ sqrt(pi T^2/3)): Baluev's bound sits
  0-15% above the simulation; a Teff without its 4 pi, a gamma-function factor
  with shifted arguments, or a (1 - z) exponent off by 1/2 move the ratio
  outside [0.95, 1.35] (measured: 0.17-0.76 or 1.6-2.6);
* the gamma-function factor against its Stirling limit sqrt(N_H/2), which is
  the reference's own use_gamma=False branch, at N = 20,000;
* the value the reference's notebook prints for its own seeded example
  (notebooks/Lomb Scargle.ipynb, cell 3: 1.09e-4), loosely (+-30%): the
  notebook's power comes from the reference's FP32 GPU path and its
  autofrequency grid, the oracle's from FP64 direct sums on the same grid.
"""
import numpy as np

from oracle import gls as og


def _noise_powers(t, freqs, ntrial, rng):
    """GLS powers (floating mean, equal weights) of ntrial pure-noise series
    on fixed times t: array (ntrial, len(freqs))."""
    N = len(t)
    w = np.full(N, 1.0 / N)
    ph = 2 * np.pi * np.multiply.outer(t, freqs)
    c, s = np.cos(ph), np.sin(ph)
    C, S = w @ c, w @ s
    C2, S2 = w @ np.cos(2 * ph), w @ np.sin(2 * ph)
    Y = rng.normal(size=(ntrial, N))
    ybar = Y @ w
    R = Y - ybar[:, None]
    YY = np.sum(w * R * R, axis=1)
    Yw = w * R
    return og.power_from_sums(C[None], S[None], C2[None], S2[None], Yw @ c, Yw @ s,
                              YY[:, None], ybar[:, None])


def test_single_frequency_tail_is_beta():
    rng = np.random.default_rng(0)
    N, ntrial = 12, 200_000
    t = np.sort(rng.uniform(0.0, 1.0, N))
    z = _noise_powers(t, np.array([3.3]), ntrial, rng)[:, 0]
    for z0 in (0.1, 0.3, 0.5, 0.7):
        emp = float(np.mean(z > z0))
        tail = float(og.fap_baluev(t, np.ones(N), np.array([z0]), 1e-12)[0])
        assert abs(tail - (1.0 - z0) ** ((N - 3) / 2.0)) < 1e-9
        sd = np.sqrt(tail * (1 - tail) / ntrial)
        assert abs(emp - tail) < 5 * sd + 1e-4, (z0, emp, tail)


def test_fap_matches_simulated_maximum_over_frequencies():
    rng = np.random.default_rng(1)
    for N, fmax in ((16, 30.0), (40, 20.0)):
        t = np.sort(rng.uniform(0.0, 1.0, N))
        df = 1.0 / (10.0 * (t.max() - t.min()))
        freqs = df * np.arange(1, int(fmax / df) + 1)
        zmax = np.concatenate([_noise_powers(t, freqs, 2000, rng).max(axis=1)
                               for _ in range(20)])
        for q in (0.1, 0.05):
            z0 = float(np.quantile(zmax, 1.0 - q))
            fap = float(og.fap_baluev(t, np.ones(N), np.array([z0]), fmax)[0])
            print(f"N={N} fmax={fmax}: simulated FAP {q} at z={z0:.4f}, Baluev {fap:.4f}")
            assert 0.95 <= fap / q <= 1.35, (N, q, fap)


def _tau(t, dy, z, fmax, use_gamma):
    """tau from the function itself: 1 - fap = Psing exp(-tau), Psing = 1 - fap(fmax -> 0)."""
    f = og.fap_baluev(t, dy, np.array([z]), fmax, use_gamma=use_gamma)[0]
    f0 = og.fap_baluev(t, dy, np.array([z]), 1e-300, use_gamma=use_gamma)[0]
    return -np.log((1.0 - f) / (1.0 - f0))


def test_gamma_factor_reaches_its_stirling_limit():
    rng = np.random.default_rng(2)
    N = 20_000
    t = np.sort(rng.uniform(0.0, 100.0, N))
    dy = rng.uniform(0.5, 1.5, N)
    z = 5e-4
    r = _tau(t, dy, z, 3.0, True) / _tau(t, dy, z, 3.0, False)
    assert abs(r - 1.0) < 1e-3, r
    # and the factor is not N-independent: at N = 10 the two differ
    t10, dy10 = t[:10], dy[:10]
    r10 = _tau(t10, dy10, 0.3, 3.0, True) / _tau(t10, dy10, 0.3, 3.0, False)
    assert abs(r10 - 1.0) > 0.02


def test_reference_notebook_example():
    """notebooks/Lomb Scargle.ipynb: data(ndata=100, freq=0.1, y0=10,
    amp=0.05, sigma=0.08, baseline=365, seed=100), autofrequency grid,
    fap_baluev(t, dy, max(powers), max(freqs)) printed as 1.09e-4."""
    rand = np.random.RandomState(100)
    t = 365.0 * np.sort(rand.rand(100))
    y = 10.0 + 0.05 * np.cos(2 * np.pi * 0.1 * t)
    dy = 0.08 * np.ones_like(y)
    y += dy * rand.randn(len(t))
    freqs = og.autofrequency(t)
    P = og.gls_direct(t, y, dy, freqs)
    fap = float(og.fap_baluev(t, dy, np.array([P.max()]), freqs.max())[0])
    print(f"notebook example: z = {P.max():.4f}, FAP = {fap:.3e} (notebook: 1.09e-4)")
    assert abs(fap / 1.09e-4 - 1.0) < 0.30
