"""The committed fixtures are the oracle's outputs on the generators' inputs
(CPU; a few entries recomputed per fixture)."""
import os

import numpy as np

from oracle import gls as og

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_served_fixture_matches_oracle_and_generator():
    from cuvarbase_amd import synthetic

    g = np.load(os.path.join(GOLDEN, "served_oracle.npz"))
    nreq = int(g["nreq"])
    _, data, freqs = synthetic.stream_requests(1.0, nreq=nreq)
    off = g["off"]
    for i in range(nreq):
        assert len(data[i][0]) == int(g["n"][i]) and len(freqs[i]) == int(g["nf"][i])
        assert float(np.sum(data[i][0])) == float(g["tsum"][i])
    for i in (0, 1, int(np.argmax(g["n"]))):
        t, y, dy = data[i]
        sel = g["sel"][off[i]:off[i + 1]]
        ref = og.gls_direct(t, y, dy, freqs[i][sel])
        assert np.max(np.abs(ref - g["p"][off[i]:off[i + 1]])) < 1e-12
        best = int(g["best"][i])
        assert best in set(sel.tolist())
        # the recorded best is a local maximum of the direct sums
        k = int(np.searchsorted(sel, best))
        for j in (k - 1, k + 1):
            if 0 <= j < len(sel) and abs(int(sel[j]) - best) == 1:
                assert ref[k] >= ref[j]
