"""Golden vectors written by the reference's own CPU path
(Code/test/test_timesC.py), for the parts of the oracle that no closed form,
brute force or NumPy routine pins:

  - the eigh-fallback trigger of _extract_shape (test_timesC.py:159-192):
    lambda <= 0 or |sum(lambda c - M c)| > 0.01 m after min(m, 1000) power
    steps from the start vector;
  - the _kshape trajectory (test_timesC.py:213-265) with the literal
    alignment-skip rule `cur_center.sum() != 0` (:128) and exactly
    min(m, 1000) power steps;
  - _preSBD_extract_shape (:127-133) and _ncc_c_3dim (:105-114) on small
    inputs.

Runs only where the reference checkout exists (REF below, this development
container); the tests read the .npz it writes and never the reference.

    python tests/golden/make_ref_fixtures.py

Caller-side shims only (the reference file is never edited):
  - np.Inf = np.inf (removed in NumPy 2, used at :85, :97, :108);
  - np.random.rand returns the seeded start vector power_v0(seed, iter,
    cluster, m) of the cluster being extracted (the reference draws an
    unseeded rand(m) at :31);
  - the module's randint returns the injected initial labels (:217).
"""
import contextlib
import importlib.util
import io
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "times-c_amd"))

import timesc  # noqa: E402  (host-side generator and power_v0; no GPU call)

REF = "/root/reference/Code/test/test_timesC.py"
OUT = os.path.join(HERE, "ref_fixtures.npz")

# small random clusters built by fallback_cluster(): the branch flag of
# every seed is kept, the centroid of the first FB_KEEP seeds and of every
# seed that takes the eigh branch
FALLBACK_SEEDS = list(range(8000))
FB_KEEP = 200


def fallback_cluster(seed):
    """(A, m): count ~ U{1..5} rows of length m ~ U{3..39}, N(0,1) values,
    rounded to integers for seeds divisible by 3 (ties and exact zeros)."""
    rng = np.random.default_rng(seed)
    m = int(rng.integers(3, 40))
    count = int(rng.integers(1, 6))
    A = rng.standard_normal((count, m))
    if seed % 3 == 0:
        A = np.round(A)
    return A.astype(np.float32).astype(np.float64), m


def load_reference():
    np.Inf = np.inf
    spec = importlib.util.spec_from_file_location("ref_test_timesC", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class StartVectors:
    """Stands in for np.random.rand: returns the vector set for the next draw."""

    def __init__(self):
        self.vec = None
        self.draws = 0

    def __call__(self, size):
        assert self.vec is not None and self.vec.size == size, (size, self.vec)
        self.draws += 1
        v, self.vec = self.vec, None
        return v.copy()


def extract_cases(mod, sv, seed_v0=0):
    """_extract_shape on each fallback cluster, start vector power_v0(seed_v0, 0, 0, m).
    Returns the branch flag of every seed, and the seeds and centroids
    (zero padded to 39 values) that are kept."""
    flags = np.zeros(len(FALLBACK_SEEDS), np.int8)
    kept, cents = [], []
    for i, s in enumerate(FALLBACK_SEEDS):
        A, m = fallback_cluster(s)
        sv.vec = timesc.power_v0(seed_v0, 0, 0, m)
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf), np.errstate(all="ignore"):
            c = mod._extract_shape(A.copy(), m)
        flags[i] = "more eigh" in buf.getvalue()
        if i < FB_KEEP or flags[i]:
            row = np.zeros(39)
            row[:m] = np.asarray(c).reshape(-1)
            kept.append(s)
            cents.append(row)
    return flags, np.array(kept, np.int32), np.array(cents)


def kshape_trajectory(mod, sv, X, k, init, seed):
    """The reference's _kshape on X (raw rows; it z-normalizes them itself)
    with injected initial labels and start vectors: labels after every
    iteration's assignment, and the final centroids."""
    n, m = X.shape
    hist = []
    calls = {"n": 0}
    orig_extract = mod._extract_shape
    orig_ncc = mod._ncc_c_3dim

    def extract(a, nn):
        it, j = divmod(calls["n"], k)
        calls["n"] += 1
        sv.vec = timesc.power_v0(seed, it, j, m)
        out = orig_extract(a, nn)
        sv.vec = None  # an empty cluster draws nothing
        return out

    def ncc3(x, y):
        cc = orig_ncc(x, y)
        hist.append((1 - cc.max(axis=2)).T.argmin(1))
        return cc

    mod._extract_shape, mod._ncc_c_3dim = extract, ncc3
    mod.randint = lambda lo, hi, size: np.asarray(init, dtype=np.int64).copy()
    try:
        with contextlib.redirect_stdout(io.StringIO()), np.errstate(all="ignore"):
            idx, cent, _ = mod._kshape(X.copy(), k)
    finally:
        mod._extract_shape, mod._ncc_c_3dim = orig_extract, orig_ncc
    assert np.array_equal(hist[-1], idx)
    return np.array(hist, np.int16), np.asarray(cent)


def main():
    mod = load_reference()
    sv = StartVectors()
    real_rand = np.random.rand
    np.random.rand = sv
    out = {}
    try:
        t = time.time()
        flags, kept, cents = extract_cases(mod, sv)
        out["fb_flags"] = flags  # seed i = FALLBACK_SEEDS[i] = i
        out["fb_kept"] = kept
        out["fb_centroids"] = cents
        print(f"_extract_shape: {len(FALLBACK_SEEDS)} clusters, {int(flags.sum())} take the eigh "
              f"branch ({time.time() - t:.1f} s)")

        # _ncc_c_3dim and _preSBD_extract_shape on small inputs
        rng = np.random.default_rng(77)
        x = rng.standard_normal((5, 16))
        y = rng.standard_normal((3, 16))
        y[2] = 0.0
        out["ncc_x"], out["ncc_y"], out["ncc_out"] = x, y, mod._ncc_c_3dim(x, y)
        members = rng.standard_normal((6, 24))
        center = rng.standard_normal(24)
        out["pre_members"], out["pre_center"] = members, center
        out["pre_out"] = mod._preSBD_extract_shape(members.copy(), center)

        # cfg0 trajectory (n=512, m=128, k=4, sigma 0.3, seed 0) and a
        # long-series one (n=137, m=2709, F=8192, k=2, classes 60/40, seed 1)
        for name, (seed, n, m, k, frac) in {
            "cfg0": (0, 512, 128, 4, None),
            "long": (1, 137, 2709, 2, [0.6, 0.4]),
        }.items():
            t = time.time()
            X, _ = timesc.generate(seed, n, m, k, 0.3, frac)
            init = timesc.init_labels(seed, n, k)
            hist, cent = kshape_trajectory(mod, sv, X.astype(np.float64), k, init, seed)
            out[f"{name}_shape"] = np.array([seed, n, m, k], np.int32)
            out[f"{name}_labels"] = hist
            out[f"{name}_centroids"] = cent
            print(f"_kshape {name}: {len(hist)} iterations ({time.time() - t:.1f} s)")
    finally:
        np.random.rand = real_rand
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
