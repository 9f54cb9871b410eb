"""Host-side logic of the drop-in API (no device calls)."""
import numpy as np
import pytest

from cuvarbase_amd import lombscargle as ls
from cuvarbase_amd import synthetic, utils
from oracle import gls as og


def test_utils_match_oracle():
    rng = np.random.default_rng(0)
    t = np.sort(rng.uniform(0, 30, 50))
    dy = rng.uniform(0.1, 1, 50)
    np.testing.assert_allclose(utils.weights(dy), og.weights(dy), rtol=1e-15)
    for kw in ({}, {"samples_per_peak": 3, "nyquist_factor": 3},
               {"minimum_frequency": 0.3, "maximum_frequency": 4.0}):
        np.testing.assert_allclose(utils.autofrequency(t, **kw), og.autofrequency(t, **kw),
                                   rtol=1e-15)
    np.testing.assert_allclose(utils.get_autofreqs(t, samples_per_peak=2, foo=1),
                               og.autofrequency(t, samples_per_peak=2))


def test_normalize_light_curves_copies():
    t = np.array([1.0, 2.0, 6.0])
    y = np.array([2.0, 2.0, 5.0])
    dy = np.ones(3)
    out = utils.normalize_light_curves([(t, y, dy)])
    np.testing.assert_allclose(out[0][0], t - t.mean())
    np.testing.assert_allclose(out[0][1], y - y.mean())
    assert out[0][2] is not dy
    assert t[0] == 1.0 and y[0] == 2.0


def test_k0_assertions():
    ls.check_k0(np.linspace(0.05, 50, 10000))
    with pytest.raises(AssertionError):
        ls.check_k0(0.01 * (3.5 + np.arange(10)))
    assert ls.get_k0(np.linspace(0.01, 40, 20000)) == 5


@pytest.mark.parametrize("prior", [None, 0.7])
def test_host_direct_sums_match_oracle(prior):
    rng = np.random.default_rng(1)
    t = np.sort(rng.uniform(0, 10, 60))
    y = np.sin(2 * np.pi * t / 1.3) + rng.normal(0, 0.3, 60)
    dy = np.full(60, 0.3)
    w, yw, ybar, YY = og.prepare(y, dy)
    freqs = np.linspace(0.1, 5, 57)
    P = ls.lomb_scargle_direct_sums(t, yw, w, freqs, YY, amplitude_priors=prior)
    s = og.direct_sums(t, w, yw, freqs)
    lam = None if prior is None else (prior ** -2, prior ** -2)
    np.testing.assert_allclose(P, og.power_from_sums(*s, YY, ybar, reg=lam), rtol=1e-9)


def test_multiharmonic_host_sums_match_oracle():
    rng = np.random.default_rng(2)
    t = np.sort(rng.uniform(0, 10, 40))
    y = rng.normal(size=40)
    w, yw, ybar, YY = og.prepare(y, np.ones(40))
    for f in (0.37, 1.9):
        a = ls.mhdirect_sums(t, yw, w, f, YY, nharms=3)
        b = og.mh_direct_sums(t, yw, w, f, nharms=3)
        for x, z in zip(a, b):
            np.testing.assert_allclose(x, z, atol=1e-13)
        assert ls.mhgls_from_sums(a, YY, ybar) == pytest.approx(og.mhgls_power(b, YY))


def test_fap_baluev_same_as_oracle():
    rng = np.random.default_rng(3)
    t = np.sort(rng.uniform(0, 100, 300))
    dy = rng.uniform(0.5, 1.0, 300)
    z = np.linspace(0, 0.2, 30)
    np.testing.assert_allclose(ls.fap_baluev(t, dy, z, 5.0), og.fap_baluev(t, dy, z, 5.0))


def test_process_construction_without_device():
    proc = ls.LombScargleAsyncProcess(sigma=2, m=8)
    assert proc.nfft_proc.sigma == 2 and proc.nfft_proc.m == 8
    assert proc.real_type is np.float32
    with pytest.raises(Exception):
        ls.LombScargleAsyncProcess(nharmonics=5)
    assert ls.LombScargleAsyncProcess(nharmonics=3).nharmonics == 3
    # autoset_m (cuvarbase/cunfft.py:378-412)
    p = ls.LombScargleAsyncProcess(autoset_m=True, sigma=4, tol=1e-8)
    assert p.nfft_proc.get_m(10000) == int(np.ceil(-np.log(0.25 * 1e-12) /
                                                   (np.pi * (1 - 1 / 7.0))))
    # window selection is validated at construction, without a device
    with pytest.raises(ValueError):
        ls.LombScargleAsyncProcess(nfft_window="kaiser")
    assert ls.LombScargleAsyncProcess(nfft_window="es", sigma=1.35, m=5).nfft_window == "es"


def test_synthetic_workloads_are_seeded():
    d1, f1 = synthetic.survey(5)
    d2, f2 = synthetic.survey(5)
    for a, b in zip(d1, d2):
        np.testing.assert_array_equal(a[0], b[0])
    for (t, y, dy), f in zip(d1, f1):
        assert 50 <= len(t) <= 5000
        assert np.all(np.diff(t) >= 0)
        df = 1 / (5 * (t.max() - t.min()))
        assert f[0] == pytest.approx(df) and f[-1] < 24.0
        # seasonal gaps: no observation in the unobservable part of a year
        assert np.all(((np.round(t) % 365) < 0.6 * 365 + 1) | (np.round(t) % 365 > 364))
    d, f = synthetic.single()
    assert len(d[0][0]) == 1000 and len(f[0]) == 10000
    arr, dd, ff = synthetic.stream_requests(100.0, nreq=20)
    assert np.all(np.diff(arr) > 0) and len(dd) == 20


def test_smoke_grid_passes_k0_check():
    # the entry point's smoke run goes through the drop-in's k0 check
    import __graft_entry__ as entry

    t, y, dy, freqs = entry.smoke_case()
    ls.check_k0(freqs)
    assert len(t) == len(y) == len(dy) and np.all(np.diff(t) >= 0)
