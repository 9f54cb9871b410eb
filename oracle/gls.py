"""FP64 restatement of the reference's GLS arithmetic (test infrastructure only).

Every function cites the reference file:line it restates.  Pins (the checks in
``tests/test_oracle_gls.py`` that a plausible transcription mistake would fail):

=========================  ====================================================
function                   pinned by
=========================  ====================================================
weights / prepare          exact-array tests against hand-written closed forms
autofrequency / k0         exact-array tests against hand-computed grids
direct_sums                brute-force per-frequency loop with ``math.cos``;
                           time-shift covariance (rotation by e^{-2 pi i f dt})
power_from_sums (all       (1) weighted least squares ``np.linalg.lstsq`` on the
modes, regularised)        model matrix (definition, docs/source/lomb.rst:6-37),
                           ridge rows for ``amplitude_prior``;
                           (2) the literal tau-rotated restatement below;
                           (3) noiseless sinusoid => P == 1; P in [0, 1];
                           (4) invariance under t-shift, y -> a*y + c, w -> c*w
power_rotated              agreement with the closed form (algebraic identity)
mh_direct_sums / mhgls     H=1 equals the closed form; H>1 equals lstsq with
                           2H+1 columns
fap_baluev                 tests/test_oracle_fap.py: the single-frequency
                           tail equals a simulated Beta(1, (N-3)/2); the whole
                           FAP within [0.95, 1.35] of the simulated maximum
                           power over a dense grid (plausible mistakes give
                           0.17-0.76 or 1.6-2.6); the gamma factor's Stirling
                           limit; the reference notebook's printed 1.09e-4
                           (+-30%); z=0 => fap=1, monotone in z
=========================  ====================================================
"""
from __future__ import annotations

import math

import numpy as np

STANDARD = 0        # cuvarbase/kernels/lomb.cu:13
FLOATING_MEAN = 1   # cuvarbase/kernels/lomb.cu:14
WINDOW = 2          # cuvarbase/kernels/lomb.cu:15


# ----------------------------------------------------------------------------
# data preparation
# ----------------------------------------------------------------------------
def weights(dy):
    """w = dy^-2 / sum(dy^-2)  (cuvarbase/utils.py:10-13)."""
    w = np.power(np.asarray(dy, dtype=np.float64), -2.0)
    return w / np.sum(w)


def normalize_light_curve(t, y, dy):
    """Subtract nanmean from t and y; dy untouched (cuvarbase/utils.py:113-146)."""
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    return t - np.nanmean(t), y - np.nanmean(y), np.asarray(dy, dtype=np.float64)


def prepare(y, dy):
    """(w, yw, ybar, YY) as LombScargleMemory.setdata (cuvarbase/lombscargle.py:247-257)."""
    w = weights(dy)
    y = np.asarray(y, dtype=np.float64)
    ybar = float(np.dot(y, w))
    yw = w * (y - ybar)
    YY = float(np.dot(w, (y - ybar) ** 2))
    return w, yw, ybar, YY


def autofrequency(t, nyquist_factor=5, samples_per_peak=5,
                  minimum_frequency=None, maximum_frequency=None):
    """Frequency grid df*(nf0 + arange(Nf))  (cuvarbase/utils.py:46-97)."""
    t = np.asarray(t, dtype=np.float64)
    baseline = t.max() - t.min()
    df = 1.0 / (baseline * samples_per_peak)
    nf0 = 1
    if minimum_frequency is not None:
        nf0 = max(nf0, int(minimum_frequency / df))
    if maximum_frequency is not None:
        nfreq = int(maximum_frequency / df) - nf0
    else:
        nfreq = int(0.5 * samples_per_peak * nyquist_factor * len(t))
    return df * (nf0 + np.arange(nfreq))


def get_k0(freqs):
    """max(1, round(f0/df))  (cuvarbase/lombscargle.py:25-26)."""
    return max(1, int(round(freqs[0] / (freqs[1] - freqs[0]))))


def check_k0(freqs, k0=None, rtol=1e-2, atol=1e-7):
    """|k0*df - f0| < rtol*df + atol  (cuvarbase/lombscargle.py:29-33)."""
    k0 = get_k0(freqs) if k0 is None else k0
    df = freqs[1] - freqs[0]
    return abs(k0 * df - freqs[0]) < rtol * df + atol


# ----------------------------------------------------------------------------
# FP64 direct sums
# ----------------------------------------------------------------------------
def direct_sums(t, w, yw, freqs, tref=0.0, chunk=256):
    """C, S, C2, S2, YC, YS at each frequency in FP64.

    C = sum w cos(phi), S = sum w sin(phi), C2/S2 at 2*phi, YC/YS with yw,
    phi = 2 pi ((t - tref) f mod 1).  The mod-1 reduction follows
    cuvarbase/lombscargle.py:370; the six sums are those of
    cuvarbase/kernels/lomb.cu:154-161 (without its FP32 +0.5 time shift, which
    leaves P unchanged).
    """
    t = np.asarray(t, dtype=np.float64) - tref
    w = np.asarray(w, dtype=np.float64)
    yw = np.asarray(yw, dtype=np.float64)
    freqs = np.asarray(freqs, dtype=np.float64)
    out = np.empty((6, len(freqs)), dtype=np.float64)
    for a in range(0, len(freqs), chunk):
        f = freqs[a:a + chunk]
        ph = np.multiply.outer(f, t)
        ph2 = 2.0 * ph
        ph -= np.floor(ph)
        ph2 -= np.floor(ph2)
        ph *= 2.0 * np.pi
        ph2 *= 2.0 * np.pi
        c = np.cos(ph)
        s = np.sin(ph)
        out[0, a:a + chunk] = c @ w
        out[1, a:a + chunk] = s @ w
        out[2, a:a + chunk] = np.cos(ph2) @ w
        out[3, a:a + chunk] = np.sin(ph2) @ w
        out[4, a:a + chunk] = c @ yw
        out[5, a:a + chunk] = s @ yw
    return out


# ----------------------------------------------------------------------------
# power from sums
# ----------------------------------------------------------------------------
def _sentinel(P):
    """NaN/Inf/negative -> -1  (cuvarbase/kernels/lomb.cu:79-80, :117-120)."""
    P = np.asarray(P, dtype=np.float64)
    bad = ~np.isfinite(P) | (P < 0)
    return np.where(bad, -1.0, P)


def power_from_sums(C, S, C2, S2, YC, YS, YY, ybar, mode=FLOATING_MEAN,
                    reg=None):
    """GLS power in closed form P = b^T A^-1 b / den.

    Restates the three modes of lspow (cuvarbase/kernels/lomb.cu:126-141) in
    the Zechmeister & Kuerster 2009 form that mhgls_from_sums uses
    (cuvarbase/lombscargle.py:497-536):
      FLOATING_MEAN: A = [[(1+C2)/2 - C^2, S2/2 - C S], [., (1-C2)/2 - S^2]],
                     b = (YC, YS), den = YY             (lomb.cu:36-83)
      STANDARD:      A = [[(1+C2)/2, S2/2], [., (1-C2)/2]],
                     b = (YC + ybar C, YS + ybar S), den = YY + ybar^2
                                                         (lomb.cu:85-123)
      WINDOW:        STANDARD with YCh=C, YSh=S, YY=0, ybar=1, no reg
                                                         (lomb.cu:136-137)
    reg = (lambda_c, lambda_s) adds to the diagonal of A
    (cuvarbase/lombscargle.py:103-118, :403-447).
    """
    C, S, C2, S2, YC, YS = (np.asarray(v, dtype=np.float64)
                            for v in (C, S, C2, S2, YC, YS))
    l0, l1 = (0.0, 0.0) if reg is None else (float(reg[0]), float(reg[1]))
    if mode == FLOATING_MEAN:
        CC = 0.5 * (1.0 + C2) - C * C + l0
        SS = 0.5 * (1.0 - C2) - S * S + l1
        CS = 0.5 * S2 - C * S
        bc, bs, den = YC, YS, YY
    elif mode == STANDARD:
        CC = 0.5 * (1.0 + C2) + l0
        SS = 0.5 * (1.0 - C2) + l1
        CS = 0.5 * S2
        bc, bs, den = YC + ybar * C, YS + ybar * S, YY + ybar * ybar
    elif mode == WINDOW:
        CC = 0.5 * (1.0 + C2)
        SS = 0.5 * (1.0 - C2)
        CS = 0.5 * S2
        bc, bs, den = 2.0 * C, 2.0 * S, 1.0
    else:
        return -np.ones_like(C)
    D = CC * SS - CS * CS
    with np.errstate(divide="ignore", invalid="ignore"):
        P = (SS * bc * bc + CC * bs * bs - 2.0 * CS * bc * bs) / (den * D)
    return _sentinel(P)


def power_rotated(C, S, C2, S2, YCh, YSh, YY, Y, mode=FLOATING_MEAN,
                  reg=None):
    """Literal FP64 restatement of lspow / lspow_flmean / lspow0.

    cuvarbase/kernels/lomb.cu:36-141 (tau-rotated form).  Used to pin
    power_from_sums by an algebraic identity.
    """
    C, S, C2, S2, YCh, YSh = (np.asarray(v, dtype=np.float64)
                              for v in (C, S, C2, S2, YCh, YSh))
    r0, r1, r2 = (0.0, 0.0, 0.0) if reg is None else (
        float(reg[0]), float(reg[1]), float(reg[2]) if len(reg) > 2 else 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        if mode == FLOATING_MEAN:
            tan2 = (S2 - 2 * S * C) / (C2 - (C * C - S * S))
            C2w = 1.0 / np.sqrt(1.0 + tan2 * tan2)
            S2w = tan2 * C2w
            Cw = np.sqrt(0.5 * (1.0 + C2w))
            Sw = np.sqrt(0.5 * (1.0 - C2w)) * np.where(S2w < 0, -1.0, 1.0)
            Csh = C * Cw + S * Sw
            Ssh = S * Cw - C * Sw
            CC = 0.5 * (1.0 + C2 * C2w + S2 * S2w) - Csh * Csh
            SS = 0.5 * (1.0 - C2 * C2w - S2 * S2w) - Ssh * Ssh
            xreg = r2 / (1.0 + r2)
            CC = CC + Csh * Csh * xreg + r0
            SS = SS + Ssh * Ssh * xreg + r1
            YC = (YCh + Y * C * xreg) * Cw + (YSh + Y * S * xreg) * Sw
            YS = (YSh + Y * S * xreg) * Cw - (YCh + Y * C * xreg) * Sw
            P = (YC * YC / CC + YS * YS / SS) / YY
        else:
            if mode == WINDOW:
                YCh, YSh, YY, Y, r0, r1 = C, S, 0.0, 1.0, 0.0, 0.0
            tan2 = S2 / C2
            C2w = 1.0 / np.sqrt(1.0 + tan2 * tan2)
            S2w = tan2 * C2w
            Cw = np.sqrt(0.5 * (1.0 + C2w))
            Sw = np.sqrt(0.5 * (1.0 - C2w)) * np.where(S2w < 0, -1.0, 1.0)
            YC = (YCh + Y * C) * Cw + (YSh + Y * S) * Sw
            YS = (YSh + Y * S) * Cw - (YCh + Y * C) * Sw
            CC = 0.5 * (1.0 + C2 * C2w + S2 * S2w) + r0
            SS = 0.5 * (1.0 - C2 * C2w - S2 * S2w) + r1
            P = (YC * YC / CC + YS * YS / SS) / (YY + Y * Y)
    return _sentinel(P)


def gls_direct(t, y, dy, freqs, mode=FLOATING_MEAN, amplitude_prior=None,
               normalize=True):
    """Full FP64 direct-sum GLS of one light curve at the exact ``freqs``.

    Pipeline of LombScargleAsyncProcess.run with use_fft=False
    (cuvarbase/lombscargle.py:923-1010, :644-659), in FP64 throughout.
    """
    if normalize:
        t, y, dy = normalize_light_curve(t, y, dy)
    w, yw, ybar, YY = prepare(y, dy)
    tref = float(np.min(t))
    s = direct_sums(t, w, yw, freqs, tref=tref)
    reg = None
    if amplitude_prior is not None:
        lam = float(amplitude_prior) ** -2
        reg = (lam, lam)
    return power_from_sums(*s, YY, ybar, mode=mode, reg=reg)


# ----------------------------------------------------------------------------
# multi-harmonic GLS (cuvarbase/lombscargle.py:337-576)
# ----------------------------------------------------------------------------
def mh_direct_sums(t, yw, w, freq, nharms=1):
    """C, S, CC, CS, SS, YC, YS for H harmonics (cuvarbase/lombscargle.py:337-400)."""
    t = np.asarray(t, dtype=np.float64)
    phase = 2 * np.pi * ((t * freq) % 1.0)
    ns = np.arange(2 * nharms + 1)
    c = np.array([np.dot(w, np.cos(n * phase)) for n in ns])
    s = np.array([np.dot(w, np.sin(n * phase)) for n in ns])
    hs = ns[1:nharms + 1]
    yc = np.array([np.dot(yw, np.cos(n * phase)) for n in hs])
    ys = np.array([np.dot(yw, np.sin(n * phase)) for n in hs])
    sgn = lambda v: 1.0 if v == 0 else float(np.sign(v))  # noqa: E731
    cc = np.array([[0.5 * (c[n + k] + c[abs(n - k)]) for k in hs] for n in hs])
    cs = np.array([[0.5 * (s[n + k] - sgn(n - k) * s[abs(n - k)]) for k in hs]
                   for n in hs])
    ss = np.array([[0.5 * (c[abs(n - k)] - c[n + k]) for k in hs] for n in hs])
    C, S = c[1:nharms + 1], s[1:nharms + 1]
    ybar = np.sum(yw)
    return (C, S, cc - np.outer(C, C), cs - np.outer(C, S),
            ss - np.outer(S, S), yc - ybar * C, ys - ybar * S)


def mhgls_power(sums, YY, amplitude_prior=None):
    """P = b^T (A + D)^-1 b / YY  (cuvarbase/lombscargle.py:403-536)."""
    C, S, CC, CS, SS, YC, YS = sums
    H = len(C)
    D = np.zeros(H) if amplitude_prior is None else (
        np.ones(H) * np.power(amplitude_prior, -2.0))
    A = np.block([[CC + np.diag(D), CS], [CS.T, SS + np.diag(D)]])
    b = np.concatenate((YC, YS))
    theta = np.linalg.solve(A, b)
    return float(theta @ b) / YY


# ----------------------------------------------------------------------------
# least-squares definition (docs/source/lomb.rst:6-37)
# ----------------------------------------------------------------------------
def lstsq_power(t, y, dy, freq, mode=FLOATING_MEAN, amplitude_prior=None,
                nharms=1):
    """P = (chi2_0 - min chi2(theta) [+ ridge]) / chi2_0 by np.linalg.lstsq."""
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    w = weights(dy)
    sw = np.sqrt(w)
    cols = []
    for h in range(1, nharms + 1):
        cols.append(np.cos(2 * np.pi * h * freq * t))
    for h in range(1, nharms + 1):
        cols.append(np.sin(2 * np.pi * h * freq * t))
    npen = len(cols)
    if mode == FLOATING_MEAN:
        cols.append(np.ones_like(t))
    X = np.stack(cols, axis=1) * sw[:, None]
    yv = y * sw
    if amplitude_prior is not None:
        lam = float(amplitude_prior) ** -2
        pen = np.zeros((npen, X.shape[1]))
        pen[np.arange(npen), np.arange(npen)] = np.sqrt(lam)
        X = np.vstack([X, pen])
        yv = np.concatenate([yv, np.zeros(npen)])
    theta, *_ = np.linalg.lstsq(X, yv, rcond=None)
    chi2 = float(np.sum((yv - X @ theta) ** 2))
    if mode == FLOATING_MEAN:
        ybar = float(np.dot(w, y))
        chi2_0 = float(np.dot(w, (y - ybar) ** 2))
    else:
        chi2_0 = float(np.dot(w, y * y))
    return (chi2_0 - chi2) / chi2_0


# ----------------------------------------------------------------------------
# significance (cuvarbase/lombscargle.py:1115-1184)
# ----------------------------------------------------------------------------
def fap_baluev(t, dy, z, fmax, d_K=3, d_H=1, use_gamma=True):
    """Baluev (2008) false-alarm probability, restated with its exponent as
    written at cuvarbase/lombscargle.py:1177 ((z/pi)^0.5 * (d-1))."""
    from scipy.special import gammaln

    t = np.asarray(t, dtype=np.float64)
    N = len(t)
    d = d_K - d_H
    N_K = N - d_K
    N_H = N - d_H
    g = (0.5 * N_H) ** (0.5 * (d - 1))
    if use_gamma:
        g = np.exp(gammaln(0.5 * N_H) - gammaln(0.5 * (N_K + 1)))
    w = np.power(np.asarray(dy, dtype=np.float64), -2)
    tbar = np.dot(w, t) / np.sum(w)
    Dt = np.dot(w, np.power(t - tbar, 2)) / np.sum(w)
    Teff = np.sqrt(4 * np.pi * Dt)
    W = fmax * Teff
    A = (2 * np.pi ** 1.5) * W
    z = np.asarray(z, dtype=np.float64)
    eZ1 = (z / np.pi) ** 0.5 * (d - 1)
    eZ2 = (1 - z) ** (0.5 * (N_K - 1))
    tau = (g * A / (2 * np.pi)) * eZ1 * eZ2
    Psing = 1 - (1 - z) ** (0.5 * N_K)
    return 1 - Psing * np.exp(-tau)


def brute_sums_scalar(t, w, yw, f):
    """Scalar pure-Python sums at one frequency (tiny inputs only)."""
    out = [0.0] * 6
    for tj, wj, ywj in zip(t, w, yw):
        ph = 2 * math.pi * tj * f
        out[0] += wj * math.cos(ph)
        out[1] += wj * math.sin(ph)
        out[2] += wj * math.cos(2 * ph)
        out[3] += wj * math.sin(2 * ph)
        out[4] += ywj * math.cos(ph)
        out[5] += ywj * math.sin(ph)
    return out
