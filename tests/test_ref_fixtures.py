"""Pins the FP64 oracle against golden vectors that the reference's own CPU
path wrote (tests/golden/ref_fixtures.npz, made by
tests/golden/make_ref_fixtures.py from Code/test/test_timesC.py):

  - the eigh-fallback trigger of _extract_shape (test_timesC.py:159-192) on
    8000 small clusters, 5 of which take the branch, and the centroids;
  - _ncc_c_3dim (:105-114) and _preSBD_extract_shape (:127-133);
  - whole _kshape trajectories (:213-265) under the reference's literal
    alignment-skip rule (:128) and exact power-step count (:152-157).
CPU only."""
import os

import numpy as np
import pytest

import oracle_py as O
import timesc as T

FIX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "ref_fixtures.npz"))


def fallback_cluster(seed):
    """The cluster of make_ref_fixtures.fallback_cluster(seed)."""
    rng = np.random.default_rng(seed)
    m = int(rng.integers(3, 40))
    count = int(rng.integers(1, 6))
    A = rng.standard_normal((count, m))
    if seed % 3 == 0:
        A = np.round(A)
    return A.astype(np.float32).astype(np.float64), m


def test_fallback_trigger_matches_reference():
    """The oracle takes the eigh branch on exactly the clusters on which the
    reference printed "more eigh" (same start vectors, exact step count)."""
    flags = FIX["fb_flags"]
    got = np.zeros_like(flags)
    for s in range(flags.size):
        A, m = fallback_cluster(s)
        _, info = O.extract_shape(A, T.power_v0(0, 0, 0, m), exact_steps=True)
        got[s] = info["fallback"]
    assert int(flags.sum()) >= 5
    assert np.array_equal(np.nonzero(got)[0], np.nonzero(flags)[0])


def test_extract_shape_centroids_match_reference():
    flags = FIX["fb_flags"]
    for s, ref in zip(FIX["fb_kept"], FIX["fb_centroids"]):
        A, m = fallback_cluster(int(s))
        got, info = O.extract_shape(A, T.power_v0(0, 0, 0, m), exact_steps=True)
        assert info["fallback"] == flags[s]
        assert np.abs(got - ref[:m]).max() <= 1e-8, (int(s), np.abs(got - ref[:m]).max())
        assert np.all(ref[m:] == 0)


def test_ncc_3dim_matches_reference():
    x, y, ref = FIX["ncc_x"], FIX["ncc_y"], FIX["ncc_out"]
    for j in range(y.shape[0]):
        for i in range(x.shape[0]):
            assert np.abs(O.ncc(x[i], y[j]) - ref[j, i]).max() <= 1e-14
    assert np.all(ref[2] == 0)  # zero-norm centroid: den = inf


def test_pre_sbd_alignment_matches_reference():
    members, center, ref = FIX["pre_members"], FIX["pre_center"], FIX["pre_out"]
    for row, want in zip(members, ref):
        s = O.align_shift(row, center, O.CONV_CPU)
        assert np.array_equal(O.roll_zeropad(row, s), want)


@pytest.mark.parametrize("name", ["cfg0", "long"])
def test_kshape_trajectory_matches_reference(name):
    """The reference's _kshape (hard-coded 10 iterations) on the seeded
    generator data with injected initial labels and start vectors: the
    oracle with the literal skip rule and the exact step count gives the
    same labels after every iteration and the same final centroids."""
    seed, n, m, k = (int(v) for v in FIX[f"{name}_shape"])
    frac = [0.6, 0.4] if name == "long" else None
    X, _ = T.generate(seed, n, m, k, 0.3, frac)
    Z = O.zscore(X.astype(np.float64))
    hist, cent = O.kshape(Z, k, init_labels=T.init_labels(seed, n, k), max_iter=10, seed=seed,
                          exact_steps=True, skip_rule=O.SKIP_NUMPY_SUM)
    ref = FIX[f"{name}_labels"]
    assert hist.shape == ref.shape and np.array_equal(hist, ref)
    assert np.abs(cent - FIX[f"{name}_centroids"]).max() <= 1e-8
