"""Write the committed FP64 fixtures from the oracle (run on the CPU).

    python tests/golden/make_golden.py

single_direct.npz   single-curve workload (seed 1): freqs and FP64 direct-sum
                    GLS power (floating mean) at the exact frequencies.
short_direct.npz    the first 8 curves of the short-curve workload (seed 5):
                    inputs and FP64 direct-sum powers on a 2,000-frequency
                    sub-grid (every 10th frequency).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cuvarbase_amd import synthetic  # noqa: E402
from oracle import gls as og  # noqa: E402


def main():
    data, freqs = synthetic.single()
    t, y, dy = data[0]
    P = og.gls_direct(t, y, dy, freqs[0])
    np.savez_compressed(os.path.join(HERE, "single_direct.npz"), freqs=freqs[0], power=P)

    data, freqs = synthetic.short(ncurves=8)
    sub = freqs[0][::10]
    out = {"freqs_sub": sub}
    for i, (t, y, dy) in enumerate(data):
        out[f"t{i}"], out[f"y{i}"], out[f"dy{i}"] = t, y, dy
        out[f"p{i}"] = og.gls_direct(t, y, dy, sub)
    np.savez_compressed(os.path.join(HERE, "short_direct.npz"), **out)
    print("wrote fixtures")


if __name__ == "__main__":
    main()
