"""Writes the golden fixtures of the parity tests: the FP64 oracle's
per-iteration labels and final centroids on the seeded benchmark inputs.

    python tests/golden/make_fixtures.py [cfg0] [cfg1]

cfg0: plain k-Shape, n=512, m=128, k=4, sigma=0.3, random-index init,
      20 iterations (seed 0).
cfg1: long series, n=1370, m=2709 (FFT length 8192), k=2, classes 60/40,
      sigma=0.3, random-centre init (init_chose=1), 20 iterations (seed 1).
The inputs are regenerated from the seed by the engine's generator; only
the oracle's outputs are stored.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "times-c_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle_py  # noqa: E402
import timesc  # noqa: E402

CONFIGS = {
    "cfg0": dict(seed=0, n=512, m=128, k=4, sigma=0.3, frac=None, init_mode=0, iters=20),
    "cfg1": dict(seed=1, n=1370, m=2709, k=2, sigma=0.3, frac=[0.6, 0.4], init_mode=1, iters=20),
}


def make(name):
    c = CONFIGS[name]
    X, y = timesc.generate(c["seed"], c["n"], c["m"], c["k"], c["sigma"], c["frac"])
    Z = oracle_py.zscore(X.astype(np.float64))
    kw = {}
    if c["init_mode"] == 0:
        kw["init_labels"] = timesc.init_labels(c["seed"], c["n"], c["k"])
    else:
        kw["init_rows"] = timesc.init_rows(c["seed"], c["n"], c["k"])
    t = time.time()
    hist, cent = oracle_py.kshape(Z, c["k"], init_mode=c["init_mode"], max_iter=c["iters"],
                                  seed=c["seed"], **kw)
    out = os.path.join(HERE, f"{name}_oracle.npz")
    np.savez_compressed(out, labels_per_iter=hist.astype(np.int16), centroids=cent,
                        **{k: np.asarray(v if v is not None else []) for k, v in c.items()})
    print(f"{name}: {len(hist)} iterations in {time.time() - t:.1f} s -> {out}")


if __name__ == "__main__":
    for name in sys.argv[1:] or list(CONFIGS):
        make(name)
