"""Binary record formats for light-curve requests and periodogram results.

The reference takes its input as in-memory ``[(t, y, dy)]`` lists
(cuvarbase/lombscargle.py:933-937) and has no file or wire format; a serving
deployment needs one, so requests and results are framed as self-describing
little-endian records that can be concatenated in a file or a pipe.

Request record ("GLSQ")::

    offset size  field
    0      4     magic  b"GLSQ"
    4      2     version (1)
    6      2     flags: bit 0 custom frequency list, bit 1 top-k result only
    8      8     request id (u64)
    16     4     n_obs (u32)
    20     4     nf (u32)
    24     8     f0 (f64)      linear grid f_k = f0 + k df (ignored with bit 0)
    32     8     df (f64)
    40     2     k (u16)       peaks wanted with bit 1
    42     1     mode (u8)     0 standard, 1 floating mean, 2 window
    43     5     reserved (0)
    48           t[n_obs], y[n_obs], dy[n_obs] (f64); then freqs[nf] (f64) with bit 0

Result record ("GLSR")::

    0      4     magic  b"GLSR"
    4      2     version (1)
    6      2     kind: 0 full periodogram, 1 top-k peaks
    8      8     request id (u64)
    16     4     status (i32, 0 = ok; a libgls status code otherwise)
    20     4     count (u32): nf for kind 0, k for kind 1
    24     8     latency (f64 seconds, arrival to completion; NaN if unknown)
    32           kind 0: power[nf] (f32)
                 kind 1: power[k] (f32), index[k] (i32), freq[k] (f64),
                         fap[k] (f64, Baluev 2008 as in fap_baluev)

Invalid powers carry the -1 sentinel of the reference kernels
(cuvarbase/kernels/lomb.cu:138-139); missing peaks have index -1.
"""
from __future__ import annotations

import struct

import numpy as np

REQ_MAGIC = b"GLSQ"
RES_MAGIC = b"GLSR"
VERSION = 1
FLAG_CUSTOM_FREQS = 1
FLAG_TOPK = 2
KIND_POWER = 0
KIND_TOPK = 1

_REQ_HEAD = struct.Struct("<4sHHQIIddHB5x")
_RES_HEAD = struct.Struct("<4sHHQiId")
assert _REQ_HEAD.size == 48 and _RES_HEAD.size == 32


class WireError(ValueError):
    """A record that does not follow the format."""


class RequestRecord:
    """Decoded request: one light curve and what to compute for it."""

    def __init__(self, rid, t, y, dy, f0=None, df=None, nf=None, freqs=None, k=0,
                 mode=1):
        self.id = int(rid)
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.dy = np.ascontiguousarray(dy, dtype=np.float64)
        if not (len(self.t) == len(self.y) == len(self.dy)):
            raise WireError("t, y and dy must have the same length")
        self.freqs = None if freqs is None else np.ascontiguousarray(freqs, np.float64)
        if self.freqs is None:
            if f0 is None or df is None or nf is None:
                raise WireError("a linear grid needs f0, df and nf")
            self.f0, self.df, self.nf = float(f0), float(df), int(nf)
        else:
            self.f0, self.df, self.nf = 0.0, 0.0, len(self.freqs)
        self.k = int(k)
        self.mode = int(mode)

    def frequencies(self):
        if self.freqs is not None:
            return self.freqs
        return self.f0 + self.df * np.arange(self.nf)


def encode_request(rec):
    flags = (FLAG_CUSTOM_FREQS if rec.freqs is not None else 0) | \
        (FLAG_TOPK if rec.k > 0 else 0)
    head = _REQ_HEAD.pack(REQ_MAGIC, VERSION, flags, rec.id, len(rec.t), rec.nf, rec.f0,
                          rec.df, rec.k, rec.mode)
    parts = [head, rec.t.tobytes(), rec.y.tobytes(), rec.dy.tobytes()]
    if rec.freqs is not None:
        parts.append(rec.freqs.tobytes())
    return b"".join(parts)


def _read_exact(stream, n):
    buf = stream.read(n)
    if buf is None:
        buf = b""
    while len(buf) < n:
        more = stream.read(n - len(buf))
        if not more:
            break
        buf += more
    return buf


def read_request(stream):
    """Next request record from a binary stream, or None at a clean EOF."""
    head = _read_exact(stream, _REQ_HEAD.size)
    if not head:
        return None
    if len(head) < _REQ_HEAD.size:
        raise WireError("truncated request header")
    magic, ver, flags, rid, n_obs, nf, f0, df, k, mode = _REQ_HEAD.unpack(head)
    if magic != REQ_MAGIC:
        raise WireError(f"bad request magic {magic!r}")
    if ver != VERSION:
        raise WireError(f"unsupported request version {ver}")
    if mode not in (0, 1, 2):
        raise WireError(f"bad mode {mode}")
    nbytes = 24 * n_obs + (8 * nf if flags & FLAG_CUSTOM_FREQS else 0)
    body = _read_exact(stream, nbytes)
    if len(body) < nbytes:
        raise WireError("truncated request body")
    arr = np.frombuffer(body, dtype="<f8")
    t, y, dy = arr[:n_obs], arr[n_obs:2 * n_obs], arr[2 * n_obs:3 * n_obs]
    freqs = arr[3 * n_obs:] if flags & FLAG_CUSTOM_FREQS else None
    return RequestRecord(rid, t, y, dy, f0=f0, df=df, nf=nf, freqs=freqs,
                         k=k if flags & FLAG_TOPK else 0, mode=mode)


def iter_requests(stream):
    while True:
        r = read_request(stream)
        if r is None:
            return
        yield r


class ResultRecord:
    def __init__(self, rid, status=0, power=None, top_power=None, top_index=None,
                 top_freq=None, fap=None, latency=float("nan")):
        self.id = int(rid)
        self.status = int(status)
        self.power = power
        self.top_power = top_power
        self.top_index = top_index
        self.top_freq = top_freq
        self.fap = fap
        self.latency = float(latency)

    @property
    def kind(self):
        return KIND_TOPK if self.top_power is not None else KIND_POWER


def encode_result(res):
    if res.kind == KIND_POWER:
        p = np.ascontiguousarray(res.power if res.power is not None else [], dtype="<f4")
        return _RES_HEAD.pack(RES_MAGIC, VERSION, KIND_POWER, res.id, res.status, len(p),
                              res.latency) + p.tobytes()
    k = len(res.top_power)
    return b"".join([
        _RES_HEAD.pack(RES_MAGIC, VERSION, KIND_TOPK, res.id, res.status, k, res.latency),
        np.ascontiguousarray(res.top_power, dtype="<f4").tobytes(),
        np.ascontiguousarray(res.top_index, dtype="<i4").tobytes(),
        np.ascontiguousarray(res.top_freq, dtype="<f8").tobytes(),
        np.ascontiguousarray(res.fap, dtype="<f8").tobytes(),
    ])


def read_result(stream):
    head = _read_exact(stream, _RES_HEAD.size)
    if not head:
        return None
    if len(head) < _RES_HEAD.size:
        raise WireError("truncated result header")
    magic, ver, kind, rid, status, count, lat = _RES_HEAD.unpack(head)
    if magic != RES_MAGIC:
        raise WireError(f"bad result magic {magic!r}")
    if ver != VERSION:
        raise WireError(f"unsupported result version {ver}")
    if kind == KIND_POWER:
        body = _read_exact(stream, 4 * count)
        if len(body) < 4 * count:
            raise WireError("truncated result body")
        return ResultRecord(rid, status, power=np.frombuffer(body, "<f4").copy(), latency=lat)
    if kind != KIND_TOPK:
        raise WireError(f"bad result kind {kind}")
    body = _read_exact(stream, 24 * count)
    if len(body) < 24 * count:
        raise WireError("truncated result body")
    p = np.frombuffer(body[:4 * count], "<f4").copy()
    i = np.frombuffer(body[4 * count:8 * count], "<i4").copy()
    f = np.frombuffer(body[8 * count:16 * count], "<f8").copy()
    a = np.frombuffer(body[16 * count:], "<f8").copy()
    return ResultRecord(rid, status, top_power=p, top_index=i, top_freq=f, fap=a,
                        latency=lat)


def iter_results(stream):
    while True:
        r = read_result(stream)
        if r is None:
            return
        yield r
