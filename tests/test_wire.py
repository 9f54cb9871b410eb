"""Request / result record formats (host only)."""
import io

import numpy as np
import pytest

from cuvarbase_amd import wire


def _req(rid=7, n=13, custom=False, k=0, mode=1):
    rng = np.random.default_rng(rid)
    t = np.sort(rng.uniform(0, 10, n))
    y = rng.normal(size=n)
    dy = rng.uniform(0.1, 0.2, n)
    if custom:
        return wire.RequestRecord(rid, t, y, dy, freqs=rng.uniform(0.1, 5, 37), k=k, mode=mode)
    return wire.RequestRecord(rid, t, y, dy, f0=0.02, df=0.01, nf=500, k=k, mode=mode)


def test_request_header_layout():
    b = wire.encode_request(_req(n=3))
    assert b[:4] == b"GLSQ" and len(b) == 48 + 24 * 3
    b = wire.encode_request(_req(n=3, custom=True, k=5))
    assert len(b) == 48 + 24 * 3 + 8 * 37
    assert int.from_bytes(b[6:8], "little") == wire.FLAG_CUSTOM_FREQS | wire.FLAG_TOPK


def test_request_round_trip_stream():
    recs = [_req(1), _req(2, custom=True), _req(3, k=5, mode=0), _req(4, n=0)]
    buf = io.BytesIO(b"".join(wire.encode_request(r) for r in recs))
    back = list(wire.iter_requests(buf))
    assert len(back) == len(recs)
    for a, b in zip(recs, back):
        assert a.id == b.id and a.k == b.k and a.mode == b.mode and a.nf == b.nf
        np.testing.assert_array_equal(a.t, b.t)
        np.testing.assert_array_equal(a.y, b.y)
        np.testing.assert_array_equal(a.dy, b.dy)
        np.testing.assert_array_equal(a.frequencies(), b.frequencies())


def test_result_round_trip():
    p = np.linspace(0, 1, 11, dtype=np.float32)
    r1 = wire.ResultRecord(5, power=p, latency=0.25)
    r2 = wire.ResultRecord(6, top_power=p[:3], top_index=np.array([4, 1, -1], np.int32),
                           top_freq=np.array([0.1, 0.2, np.nan]), fap=np.array([1e-3, 0.5, np.nan]))
    buf = io.BytesIO(wire.encode_result(r1) + wire.encode_result(r2))
    a, b = list(wire.iter_results(buf))
    assert a.kind == wire.KIND_POWER and a.id == 5 and a.latency == 0.25
    np.testing.assert_array_equal(a.power, p)
    assert b.kind == wire.KIND_TOPK and b.id == 6 and np.isnan(b.latency)
    np.testing.assert_array_equal(b.top_index, [4, 1, -1])
    np.testing.assert_array_equal(b.top_power, p[:3])
    np.testing.assert_array_equal(b.fap[:2], [1e-3, 0.5])


def test_malformed_records_raise():
    good = wire.encode_request(_req())
    with pytest.raises(wire.WireError):
        wire.read_request(io.BytesIO(good[:-1]))
    with pytest.raises(wire.WireError):
        wire.read_request(io.BytesIO(good[:20]))
    with pytest.raises(wire.WireError):
        wire.read_request(io.BytesIO(b"XXXX" + good[4:]))
    bad_ver = good[:4] + (2).to_bytes(2, "little") + good[6:]
    with pytest.raises(wire.WireError):
        wire.read_request(io.BytesIO(bad_ver))
    with pytest.raises(wire.WireError):
        wire.RequestRecord(1, [0.0, 1.0], [1.0], [1.0, 1.0], f0=0, df=1, nf=3)
    assert wire.read_request(io.BytesIO(b"")) is None
