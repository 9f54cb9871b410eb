"""Pins of oracle/nfft.py (FP64 adjoint NFFT) against brute force and the DFT."""
import numpy as np
import pytest

from oracle import gls as og
from oracle import nfft as on


def _data(seed=100, n=100):
    rng = np.random.RandomState(seed)
    t = np.sort(rng.rand(n))
    y = np.cos(2 * np.pi * (3.0 / (t.max() - t.min())) * t) + 0.1 * rng.randn(n)
    return t, y


def test_gridding_matches_bruteforce_cell_loop():
    t, y = _data(n=20)
    g = on.gaussian_grid_reference(t, y, nf=16, sigma=2, m=4, samples_per_peak=1.0)
    b = on.gaussian_grid_bruteforce(t, y, nf=16, sigma=2, m=4, samples_per_peak=1.0)
    np.testing.assert_allclose(g, b, atol=1e-14)
    g = on.gaussian_grid_reference(t, y, nf=25, sigma=3, m=6, samples_per_peak=3.0)
    b = on.gaussian_grid_bruteforce(t, y, nf=25, sigma=3, m=6, samples_per_peak=3.0)
    np.testing.assert_allclose(g, b, atol=1e-14)


def test_tap_weights_equal_fast_gridding_recurrence():
    """psi(m + fr - k) == q1 * q2^k * q3[k] (the factorisation of
    cuvarbase/tests/test_nfft.py:61-68 used by fast_gaussian_grid)."""
    m, sigma = 8, 2.0
    b = on.gauss_b(sigma, m)
    for fr in (0.0, 0.37, 0.999):
        xg = m + fr
        q1 = np.exp(-xg ** 2 / b) / np.sqrt(np.pi * b)
        q2 = np.exp(2 * xg / b)
        q3 = np.exp(-np.arange(2 * m + 1) ** 2 / b)
        k = np.arange(2 * m + 1)
        fast = q1 * q2 ** k * q3
        direct = np.exp(-(m + fr - k) ** 2 / b) / np.sqrt(np.pi * b)
        np.testing.assert_allclose(fast, direct, rtol=1e-12)


def test_fft_convention_is_unnormalised_plus_sign():
    rng = np.random.default_rng(3)
    x = rng.normal(size=48) + 1j * rng.normal(size=48)
    n = len(x)
    dft = np.array([np.sum(x * np.exp(2j * np.pi * np.arange(n) * k / n)) for k in range(n)])
    np.testing.assert_allclose(np.fft.ifft(x) * n, dft, atol=1e-11)


@pytest.mark.parametrize("f0_over_df", [0.0, 10.0, 10.37, -250.5, 1234.25])
def test_centred_nfft_matches_direct_dft(f0_over_df):
    t, y = _data()
    t = t * 37.0 + 5.0
    df = 1.0 / (5 * (t.max() - t.min()))
    K = 500
    f0 = f0_over_df * df
    V = on.nfft_adjoint(t, y, f0, df, K, sigma=2.0, m=8)
    D = on.direct_dft(t, y, f0 + df * np.arange(K))
    # alias + truncation bound at sigma=2, m=8: ~1e-7 relative to sum|y|
    assert np.max(np.abs(V - D)) < 1e-6 * np.sum(np.abs(y))


def test_centred_nfft_rejects_dropped_phase():
    """apply_phase=False is the DFT with times measured from tref."""
    t, y = _data()
    df = 0.2
    V = on.nfft_adjoint(t + 100.0, y, 3.0, df, 200, sigma=2.0, m=8, tref=100.0,
                        apply_phase=False)
    D = on.direct_dft(t, y, 3.0 + df * np.arange(200))
    assert np.max(np.abs(V - D)) < 1e-6 * np.sum(np.abs(y))


def test_reference_layout_matches_dft_at_reference_test_sigma():
    """The reference's one-sided layout at its tests' sigma=5
    (cuvarbase/tests/test_nfft.py:24, :202-239), scaled times, integer k0."""
    t, y = _data()
    nf = 500
    tsc = (t - t.min()) / (t.max() - t.min()) - 0.5
    for k0 in (-250, 0, 10):
        G = on.nfft_adjoint_reference_style(tsc, y, nf, minimum_frequency=k0,
                                            samples_per_peak=1.0, sigma=5, m=8)
        D = on.direct_dft(tsc, y, k0 + np.arange(nf, dtype=float))
        np.testing.assert_allclose(G, D, rtol=5e-3, atol=5e-3)


def test_reference_layout_aliases_at_sigma2():
    """One-sided band at sigma=2 is at the aliasing limit: error >> centred band."""
    t, y = _data()
    nf = 500
    tsc = (t - t.min()) / (t.max() - t.min())
    G = on.nfft_adjoint_reference_style(tsc, y, nf, samples_per_peak=1.0, sigma=2, m=8)
    D = on.direct_dft(tsc, y, np.arange(nf, dtype=float))
    V = on.nfft_adjoint(tsc, y, 0.0, 1.0, nf, sigma=2.0, m=8)
    assert np.max(np.abs(G - D)) > 1e3 * np.max(np.abs(V - D))


def test_gls_nfft_matches_direct_sums_single_workload():
    from cuvarbase_amd import synthetic

    data, freqs = synthetic.single()
    t, y, dy = data[0]
    sel = np.arange(0, 10000, 37)
    Pd = og.gls_direct(t, y, dy, freqs[0][sel])
    for sigma in (2.0, 1.5):
        Pn = on.gls_nfft(t, y, dy, freqs[0], sigma=sigma, m=8)
        assert np.max(np.abs(Pn[sel] - Pd)) < (1e-6 if sigma == 2.0 else 1e-4)
    for mode in (og.STANDARD, og.WINDOW):
        Pd = og.gls_direct(t, y, dy, freqs[0][sel], mode=mode)
        Pn = on.gls_nfft(t, y, dy, freqs[0], sigma=2.0, m=8, mode=mode)
        np.testing.assert_allclose(Pn[sel], Pd, atol=1e-6)
