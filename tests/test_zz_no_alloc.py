"""The allocation-free execute (gls_lomb_execute_static) against gls_lomb_execute.

Both run the same plan on the same device inputs; the static entry point uses
the library's static device scratch instead of a caller workspace.  Two calls
in a row through the static entry point, the second with a smaller batch, each
compared with the workspace entry point: |dP| <= 2e-5 (the default spreading
adds in LDS with float atomics, so two executes of one plan may differ by
round-off) and the same best index per curve; both also against the FP64
oracle at |dP| <= 1e-4.
"""
import numpy as np
import pytest

from oracle import gls as og

pytestmark = pytest.mark.gpu

TOL = 1e-4
SAME_TOL = 2e-5


def _static(batch, out, stream):
    from cuvarbase_amd import _lib

    _lib.call("gls_lomb_execute_static", batch.plan.handle, batch.d_t.ptr, batch.d_a.ptr,
              batch.d_b.ptr, out.ptr, stream.handle)


def _compare(data, freqs, **kw):
    from cuvarbase_amd import _device, engine

    b = engine.LombBatch(data, freqs, **kw)
    s = _device.Stream()
    b.upload(stream=s)
    assert b.plan.workspace_bytes <= (128 << 20)
    out = _device.DeviceArray(max(b.total_out, 1), np.float32)
    _static(b, out, s)
    b.execute(stream=s)
    ref = b.split(b.download_async(stream=s))
    got = b.split(out.get_async(stream=s, count=b.total_out))
    s.synchronize()
    worst = 0.0
    for i, (g, r) in enumerate(zip(got, ref)):
        g = np.asarray(g, np.float64)
        r = np.asarray(r, np.float64)
        worst = max(worst, float(np.max(np.abs(g - r))))
        assert np.argmax(g) == np.argmax(r), f"curve {i}: best index differs"
    return got, worst


def _oracle_check(data, freqs, got, nspot=400, seed=0):
    rng = np.random.default_rng(seed)
    for (t, y, dy), f, p in zip(data, freqs, got):
        sel = np.unique(np.concatenate([np.arange(min(20, len(f))),
                                        rng.choice(len(f), size=min(nspot, len(f)),
                                                   replace=False),
                                        [int(np.argmax(p))]]))
        ref = og.gls_direct(t, y, dy, f[sel])
        err = float(np.max(np.abs(np.asarray(p, np.float64)[sel] - ref)))
        assert err <= TOL, err


def test_static_entry_matches_workspace_entry_twice():
    from cuvarbase_amd import synthetic, tuning

    cfg = tuning.choose_nfft_config(TOL)
    # call 1: up to 12 survey curves of fewer than 100,000 frequencies and one
    # curve of 30,000 points (chunked bucketing); call 2: a smaller batch, the
    # single 1,000-point curve
    data, freqs = synthetic.survey(60, seed=11)
    keep = [i for i, f in enumerate(freqs) if len(f) < 100_000][:12]
    data, freqs = [data[i] for i in keep], [freqs[i] for i in keep]
    rng = np.random.default_rng(12)
    t = np.sort(rng.uniform(0.0, 400.0, 30_000))
    y = np.sin(2 * np.pi * t / 3.3) + rng.normal(0.0, 0.3, t.size)
    data.append((t, y, np.full(t.size, 0.3)))
    freqs.append(synthetic.survey_grid(t, fmax=10.0))
    print(f"call 1: {len(data)} curves, {sum(len(d[0]) for d in data)} points, "
          f"{sum(len(f) for f in freqs)} frequencies", flush=True)
    got1, d1 = _compare(data, freqs, **cfg)
    print(f"call 1: max|dP| static vs workspace = {d1:.2e}", flush=True)
    assert d1 <= SAME_TOL
    _oracle_check(data[:6] + data[-1:], freqs[:6] + freqs[-1:], got1[:6] + got1[-1:])

    sd, sf = synthetic.single()
    got2, d2 = _compare(sd, sf, sigma=2.0, m=8)
    print(f"call 2 (smaller): max|dP| static vs workspace = {d2:.2e}", flush=True)
    assert d2 <= SAME_TOL
    _oracle_check(sd, sf, got2)


def test_static_entry_rejects_oversized_plan():
    from cuvarbase_amd import _device, _lib, engine, synthetic

    # a plan whose workspace exceeds the scratch: rejected on the host before
    # any launch, with a message naming both sizes
    data, freqs = synthetic.survey(60, seed=13)
    b = engine.LombBatch(data, freqs, sigma=2.0, m=8)
    assert b.plan.workspace_bytes > _lib.STATIC_SCRATCH_BYTES
    b.upload(stream=None)
    out = _device.DeviceArray(max(b.total_out, 1), np.float32)
    with pytest.raises(_lib.GLSError) as ei:
        _lib.call("gls_lomb_execute_static", b.plan.handle, b.d_t.ptr, b.d_a.ptr, b.d_b.ptr,
                  out.ptr, None)
    assert ei.value.code == _lib.GLS_E_UNSUPPORTED
    assert "static scratch" in str(ei.value)
