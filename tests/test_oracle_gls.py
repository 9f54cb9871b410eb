"""Pins of oracle/gls.py (FP64 GLS arithmetic) against independent mathematics."""
import math

import numpy as np
import pytest

from oracle import gls as og


def _curve(seed=0, n=80, T=10.0, noise=0.3):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0, T, n))
    y = 1.3 + 0.7 * np.sin(2 * np.pi * t / 1.7 + 0.4) + rng.normal(0, noise, n)
    dy = rng.uniform(0.1, 0.4, n)
    return t, y, dy


def test_weights_and_prepare_closed_form():
    y = np.array([1.0, 2.0, 4.0])
    dy = np.array([1.0, 2.0, 1.0])
    w = og.weights(dy)
    np.testing.assert_allclose(w, np.array([1.0, 0.25, 1.0]) / 2.25, rtol=0, atol=1e-15)
    w2, yw, ybar, YY = og.prepare(y, dy)
    ybar_exact = (1.0 + 0.5 + 4.0) / 2.25
    assert ybar == pytest.approx(ybar_exact, abs=1e-14)
    np.testing.assert_allclose(yw, w * (y - ybar_exact), atol=1e-15)
    assert YY == pytest.approx(np.sum(w * (y - ybar_exact) ** 2), abs=1e-14)
    assert abs(np.sum(yw)) < 1e-14


def test_normalize_light_curve_nanmean():
    t = np.array([1.0, 2.0, 3.0])
    y = np.array([1.0, np.nan, 5.0])
    tc, yc, _ = og.normalize_light_curve(t, y, np.ones(3))
    np.testing.assert_array_equal(tc, [-1.0, 0.0, 1.0])
    np.testing.assert_array_equal(yc[[0, 2]], [-2.0, 2.0])
    assert np.isnan(yc[1])


def test_autofrequency_hand_computed():
    t = np.array([0.0, 4.0, 10.0])
    f = og.autofrequency(t, nyquist_factor=5, samples_per_peak=5)
    df = 1.0 / 50.0
    np.testing.assert_allclose(f, df * (1 + np.arange(int(0.5 * 5 * 5 * 3))), rtol=1e-15)
    f = og.autofrequency(t, samples_per_peak=5, minimum_frequency=0.1, maximum_frequency=1.0)
    nf0 = int(0.1 / df)
    np.testing.assert_allclose(f, df * (nf0 + np.arange(int(1.0 / df) - nf0)), rtol=1e-15)


def test_k0_rules():
    freqs = np.linspace(0.05, 50.0, 10000)
    assert og.get_k0(freqs) == 10
    assert og.check_k0(freqs)
    df = 0.01
    assert og.get_k0(df * (3 + np.arange(10))) == 3
    assert not og.check_k0(df * (3.5 + np.arange(10)))
    assert og.get_k0(df * (0.2 + np.arange(10))) == 1


def test_direct_sums_match_scalar_loop():
    t, y, dy = _curve(1, n=15)
    w, yw, _, _ = og.prepare(y, dy)
    for f in (0.0371, 1.0 / 1.7, 13.25):
        s = og.direct_sums(t, w, yw, [f])[:, 0]
        ref = og.brute_sums_scalar(t, w, yw, f)
        np.testing.assert_allclose(s, ref, atol=1e-13)


def test_direct_sums_time_shift_rotation():
    t, y, dy = _curve(2)
    w, yw, _, _ = og.prepare(y, dy)
    f = np.array([0.31, 2.2, 7.9])
    dt = 123.456
    a = og.direct_sums(t, w, yw, f)
    b = og.direct_sums(t + dt, w, yw, f)
    rot1 = np.exp(2j * np.pi * f * dt)
    rot2 = np.exp(2j * np.pi * 2 * f * dt)
    np.testing.assert_allclose(b[0] + 1j * b[1], rot1 * (a[0] + 1j * a[1]), atol=1e-10)
    np.testing.assert_allclose(b[2] + 1j * b[3], rot2 * (a[2] + 1j * a[3]), atol=1e-10)
    np.testing.assert_allclose(b[4] + 1j * b[5], rot1 * (a[4] + 1j * a[5]), atol=1e-10)


@pytest.mark.parametrize("mode", [og.FLOATING_MEAN, og.STANDARD])
@pytest.mark.parametrize("prior", [None, 0.5])
def test_power_matches_weighted_least_squares(mode, prior):
    t, y, dy = _curve(3)
    tc, yc, dyc = og.normalize_light_curve(t, y, dy)
    w, yw, ybar, YY = og.prepare(yc, dyc)
    freqs = np.array([0.21, 1.0 / 1.7, 3.3, 7.77])
    s = og.direct_sums(tc, w, yw, freqs)
    lam = None if prior is None else (prior ** -2, prior ** -2)
    P = og.power_from_sums(*s, YY, ybar, mode=mode, reg=lam)
    ref = [og.lstsq_power(tc, yc, dyc, f, mode=mode, amplitude_prior=prior) for f in freqs]
    np.testing.assert_allclose(P, ref, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("mode", [og.FLOATING_MEAN, og.STANDARD, og.WINDOW])
def test_closed_form_equals_rotated_form(mode):
    t, y, dy = _curve(4)
    w, yw, ybar, YY = og.prepare(y, dy)
    freqs = np.linspace(0.05, 9.0, 301)
    s = og.direct_sums(t, w, yw, freqs)
    for reg in (None, (3.0, 3.0)):
        if mode == og.WINDOW and reg is not None:
            continue
        a = og.power_from_sums(*s, YY, ybar, mode=mode, reg=reg)
        b = og.power_rotated(*s, YY, ybar, mode=mode, reg=reg)
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-11)


def test_window_mode_literal_value():
    # one observation at t=0: C=1, S=0, C2=1, S2=0 -> A singular -> sentinel
    P = og.power_from_sums(1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, mode=og.WINDOW)
    assert P == -1
    # four points a quarter period apart: C = S = C2 = S2 = 0 -> window power 0
    P = og.power_from_sums(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, mode=og.WINDOW)
    assert P == pytest.approx(0.0)
    # window power of a single tone at C = 0.5: b = (1, 0), A = diag(3/4, 1/4) -> 4/3
    P = og.power_from_sums(0.5, 0.0, 0.5, 0.0, 0.0, 0.0, 1.0, 0.0, mode=og.WINDOW)
    assert P == pytest.approx(1.0 / 0.75)


def test_noiseless_sinusoid_gives_unit_power():
    rng = np.random.default_rng(5)
    t = np.sort(rng.uniform(0, 20, 60))
    f = 0.8
    y = 3.0 + 0.9 * np.cos(2 * np.pi * f * t) - 0.4 * np.sin(2 * np.pi * f * t)
    dy = rng.uniform(0.5, 1.5, 60)
    P = og.gls_direct(t, y, dy, [f, 1.3])
    assert P[0] == pytest.approx(1.0, abs=1e-12)
    assert 0.0 <= P[1] < 1.0
    y0 = 0.9 * np.cos(2 * np.pi * f * t) - 0.4 * np.sin(2 * np.pi * f * t)
    P0 = og.gls_direct(t, y0, dy, [f], mode=og.STANDARD, normalize=False)
    assert P0[0] == pytest.approx(1.0, abs=1e-12)


def test_power_bounds_and_invariances():
    t, y, dy = _curve(6)
    freqs = np.linspace(0.05, 12.0, 400)
    P = og.gls_direct(t, y, dy, freqs)
    assert np.all((P >= 0) & (P <= 1))
    np.testing.assert_allclose(og.gls_direct(t + 1e4, y, dy, freqs), P, atol=1e-10)
    np.testing.assert_allclose(og.gls_direct(t, -2.5 * y + 7.0, dy, freqs), P, atol=1e-10)
    np.testing.assert_allclose(og.gls_direct(t, y, 3.0 * dy, freqs), P, atol=1e-10)
    Ps = og.gls_direct(t, y, dy, freqs, mode=og.STANDARD, normalize=False)
    np.testing.assert_allclose(og.gls_direct(t + 50.0, y, dy, freqs, mode=og.STANDARD,
                                             normalize=False), Ps, atol=1e-10)


def test_constant_signal_gives_sentinel():
    t = np.linspace(0, 5, 30)
    P = og.gls_direct(t, np.full(30, 2.0), np.ones(30), [0.3, 1.1])
    np.testing.assert_array_equal(P, [-1.0, -1.0])


def test_multiharmonic_h1_equals_closed_form_and_h2_equals_lstsq():
    t, y, dy = _curve(7)
    tc, yc, dyc = og.normalize_light_curve(t, y, dy)
    w, yw, ybar, YY = og.prepare(yc, dyc)
    for f in (0.4, 1.0 / 1.7, 2.9):
        s1 = og.mh_direct_sums(tc, yw, w, f, nharms=1)
        sums = og.direct_sums(tc, w, yw, [f])
        ref = og.power_from_sums(*sums, YY, ybar)[0]
        assert og.mhgls_power(s1, YY) == pytest.approx(ref, rel=1e-10)
        s2 = og.mh_direct_sums(tc, yw, w, f, nharms=2)
        assert og.mhgls_power(s2, YY) == pytest.approx(
            og.lstsq_power(tc, yc, dyc, f, nharms=2), rel=1e-9)


def test_fap_baluev_special_values():
    t, y, dy = _curve(8, n=200)
    assert og.fap_baluev(t, dy, 0.0, 10.0) == pytest.approx(1.0)
    z = np.linspace(0.01, 0.5, 50)
    fap = og.fap_baluev(t, dy, z, 10.0)
    assert np.all(np.diff(fap) <= 1e-15)
    assert np.all((fap >= 0) & (fap <= 1))


def test_golden_single_curve_periodogram():
    """The single-curve workload (N=1000 on [0,10] d, 10^4 frequencies) against
    the committed FP64 vectors."""
    import os

    from cuvarbase_amd import synthetic

    path = os.path.join(os.path.dirname(__file__), "golden", "single_direct.npz")
    g = np.load(path)
    data, freqs = synthetic.single()
    t, y, dy = data[0]
    np.testing.assert_array_equal(freqs[0], g["freqs"])
    sel = np.arange(0, 10000, 97)
    P = og.gls_direct(t, y, dy, freqs[0][sel])
    np.testing.assert_allclose(P, g["power"][sel], rtol=0, atol=1e-12)
    assert math.isclose(1 / g["freqs"][np.argmax(g["power"])], 0.73, rel_tol=2e-3)


@pytest.mark.parametrize("nharms,prior", [(3, None), (2, 0.7), (4, 2.0)])
def test_multiharmonic_power_equals_least_squares(nharms, prior):
    """mhgls_power (cuvarbase/lombscargle.py:337-536 restated) equals the
    weighted least-squares fit of H harmonics + offset, with ridge rows for
    the amplitude prior."""
    rng = np.random.default_rng(31 + nharms)
    t = np.sort(rng.uniform(0, 20, 120))
    y = 1.0 + np.sin(2 * np.pi * t / 2.3) + 0.4 * np.cos(4 * np.pi * t / 2.3) \
        + rng.normal(0, 0.3, 120)
    dy = rng.uniform(0.2, 0.5, 120)
    tc, yc, dyc = og.normalize_light_curve(t, y, dy)
    w, yw, ybar, YY = og.prepare(yc, dyc)
    for f in (0.17, 1 / 2.3, 1.9):
        s = og.mh_direct_sums(tc, yw, w, f, nharms=nharms)
        got = og.mhgls_power(s, YY, amplitude_prior=prior)
        ref = og.lstsq_power(tc, yc, dyc, f, nharms=nharms, amplitude_prior=prior)
        assert got == pytest.approx(ref, rel=1e-8, abs=1e-12)
