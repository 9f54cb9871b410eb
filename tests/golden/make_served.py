"""Oracle values for the served-mix parity test (run on the CPU).

    python tests/golden/make_served.py

served_oracle.npz holds, for the first NREQ requests of the serving workload
(``synthetic.stream_requests``, seed 4: N log-uniform in [100, 50,000],
baselines U[30, 3650] d, df = 1/(5 T), fmax = 24 /d):

  n, nf, tsum     per request: size, grid length and sum of t (ties the
                  fixture to the generator's output)
  sel, p, off     frequency indices and their FP64 direct-sum GLS powers
                  (oracle.gls.gls_direct) of request i at [off[i], off[i+1]):
                  the 20 lowest frequencies, +-10 bins around the oracle's best
                  frequency and 60 seeded random bins
  best            per request: index of the highest power of the FP64 oracle
                  NFFT (oracle.nfft.gls_nfft, Gaussian sigma = 2, m = 12) over
                  the whole grid
  ties, toff      every index whose oracle NFFT power is within 1e-4 of the
                  best, request i at [toff[i], toff[i+1]) (a GPU best index
                  among them counts as the same best period)

The oracle NFFT is checked against the direct sums at every selected index
before anything is written (|dP| < 1e-7).
"""
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

NREQ = 200
TIE = 1e-4


def one(args):
    from oracle import gls as og
    from oracle import nfft as on

    i, (t, y, dy), f = args
    Pn = on.gls_nfft(t, y, dy, f, sigma=2.0, m=12)
    best = int(np.argmax(Pn))
    rng = np.random.default_rng(1000 + i)
    sel = np.unique(np.concatenate([
        np.arange(min(20, len(f))),
        np.arange(max(best - 10, 0), min(best + 11, len(f))),
        rng.choice(len(f), size=min(60, len(f)), replace=False)]))
    Pd = og.gls_direct(t, y, dy, f[sel])
    agree = float(np.max(np.abs(Pn[sel] - Pd)))
    ties = np.flatnonzero(Pn >= Pn[best] - TIE)
    return i, len(t), len(f), float(np.sum(t)), sel, Pd, best, ties, agree


def main():
    from cuvarbase_amd import synthetic

    _, data, freqs = synthetic.stream_requests(1.0, nreq=NREQ)
    with mp.get_context("spawn").Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(one, [(i, data[i], freqs[i]) for i in range(NREQ)], chunksize=1)
    worst = max(r[-1] for r in res)
    out = {
        "nreq": np.array(NREQ), "tie": np.array(TIE),
        "n": np.array([r[1] for r in res], dtype=np.int64),
        "nf": np.array([r[2] for r in res], dtype=np.int64),
        "tsum": np.array([r[3] for r in res], dtype=np.float64),
        "sel": np.concatenate([r[4] for r in res]).astype(np.int32),
        "p": np.concatenate([r[5] for r in res]),
        "off": np.cumsum([0] + [len(r[4]) for r in res]).astype(np.int64),
        "best": np.array([r[6] for r in res], dtype=np.int64),
        "ties": np.concatenate([r[7] for r in res]).astype(np.int32),
        "toff": np.cumsum([0] + [len(r[7]) for r in res]).astype(np.int64),
    }
    print(f"oracle NFFT vs direct sums at the selected bins: max |dP| = {worst:.2e}")
    assert worst < 1e-7
    np.savez_compressed(os.path.join(HERE, "served_oracle.npz"), **out)
    print(f"wrote served_oracle.npz ({NREQ} requests)")


if __name__ == "__main__":
    main()
