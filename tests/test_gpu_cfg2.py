"""The large single-job workload at its own size: n=100000 series of length
m=512 (FFT length 1024), k=16, generator seed 2 with noise 0.5, 10
iterations from the seeded initial labels, stopping when the labels repeat
(test_timesC.py:213-265).

Checked on every iteration of the fit:
  - the in-fit pruned assignment (labels, SBD distances, lags) equals the
    dense assignment of all 100000 rows against the same centroids, up to
    ties (SBD within 1e-6 between the two labels, NCC within 1e-6 between
    the two lags), distances within 2e-6;
  - SBD distances within 2e-6 of the FP64 oracle's on a 2000-row sample.
Iterations 1 and 2 are also matched to the FP64 oracle step by step: the
refinement from the GPU's previous labels and centroids (centroids within
1e-4 relative L2) and the assignment against the GPU's centroids (labels up
to SBD ties of 1e-6) over all 100000 rows."""
import numpy as np
import pytest

import oracle_py as O
import timesc as T
from parity import CENT_RTOL, SBD_TIE, assert_labels_match, centroid_errors

pytestmark = pytest.mark.gpu

N, M, K, SEED, SIGMA, ITERS = 100000, 512, 16, 2, 0.5, 10
DIST_TOL = 2e-6
SAMPLE = 2000


@pytest.fixture(scope="module")
def cfg2_fit(ctx):
    """Runs the fit step by step; per iteration the in-fit state and the dense
    assignment of all rows against the same centroids (on a context of its
    own, so the fit's workspace is never touched)."""
    print(f"\ncfg2: n={N} m={M} k={K} seed={SEED}", flush=True)
    X, _ = T.generate(SEED, N, M, K, SIGMA)
    d = ctx.to_device(X)
    steps = []
    other = T.Context(0)
    try:
        ctx.znorm(d, ddof=0)
        Z = d.numpy()
        init = T.init_labels(SEED, N, K)
        ctx.fit_begin(d, K, max_iter=ITERS, seed=SEED)
        dc = ctx.to_device(np.zeros((K, M), np.float32))
        try:
            for it in range(ITERS):
                changed = ctx.fit_step()
                lab, dist, sh, C = ctx.fit_snapshot(N, K, M)
                ctx.h2d(dc, C)
                dl, dd, ds, dm = other.assign(d, dc, dmat=True)
                steps.append(dict(lab=lab, dist=dist, shift=sh, C=C, dlab=dl, ddist=dd,
                                  dshift=ds, dmat=dm))
                print(f"cfg2 iteration {it + 1}: changed={changed}", flush=True)
                if not changed:
                    break
        finally:
            dc.free()
        res = ctx.fit_end(N, K, M)
    finally:
        d.free()
        other.close()
    return dict(Z=Z, init=init, steps=steps, res=res)


def test_cfg2_runs_the_configured_iterations(cfg2_fit):
    res = cfg2_fit["res"]
    assert 1 <= res["stats"]["iterations"] <= ITERS
    assert res["stats"]["iterations"] == len(cfg2_fit["steps"])
    assert np.array_equal(res["labels"], cfg2_fit["steps"][-1]["lab"])


def test_cfg2_pruned_equals_dense_every_iteration(ctx, cfg2_fit):
    Z = cfg2_fit["Z"]
    for it, s in enumerate(cfg2_fit["steps"]):
        lab, dl, dm = s["lab"], s["dlab"], s["dmat"]
        bad = np.nonzero(lab != dl)[0]
        if bad.size:
            gap = np.abs(dm[bad, lab[bad]] - dm[bad, dl[bad]])
            assert gap.max() <= SBD_TIE, (it, bad.size, float(gap.max()))
        same = lab == dl
        assert np.abs(s["dist"][same] - s["ddist"][same]).max() <= DIST_TOL, it
        # lags: equal, or the two lags' NCC values tie
        lagdiff = np.nonzero(same & (s["shift"] != s["dshift"]))[0]
        if lagdiff.size:
            rows = lagdiff[:256]
            dx = ctx.to_device(np.ascontiguousarray(Z[rows]))
            dc = ctx.to_device(np.ascontiguousarray(s["C"]))
            try:
                cc = ctx.ncc(dx, dc)  # [rows][k][2m-1]
            finally:
                dx.free()
                dc.free()
            for q, i in enumerate(rows):
                j = lab[i]
                a, b = cc[q, j, s["shift"][i] + M - 1], cc[q, j, s["dshift"][i] + M - 1]
                assert abs(a - b) <= SBD_TIE, (it, int(i), float(a), float(b))
        print(f"iteration {it + 1}: {bad.size} label ties, {lagdiff.size} lag ties", flush=True)


def test_cfg2_sbd_error_against_fp64_oracle(cfg2_fit):
    Z = cfg2_fit["Z"]
    rows = np.random.default_rng(0).choice(N, SAMPLE, replace=False)
    Zs = Z[rows].astype(np.float64)
    worst = 0.0
    for it, s in enumerate(cfg2_fit["steps"]):
        _, _, _, odm = O.assign(Zs, s["C"].astype(np.float64), dmat=True)
        err = np.abs(s["dist"][rows] - odm[np.arange(SAMPLE), s["lab"][rows]])
        worst = max(worst, float(err.max()))
        # the GPU's label is the oracle's argmin up to ties
        best = odm.min(axis=1)
        assert np.all(odm[np.arange(SAMPLE), s["lab"][rows]] - best <= SBD_TIE), it
    print(f"max SBD error vs FP64 oracle over {len(cfg2_fit['steps'])} iterations: {worst:.3e}")
    assert worst <= DIST_TOL


@pytest.mark.parametrize("it", [0, 1])
def test_cfg2_iteration_matches_oracle_stepwise(cfg2_fit, it):
    """The refinement of iteration it+1 from the GPU's previous labels,
    centroids and alignment lags (the previous assignment's), and the
    assignment against the GPU's new centroids.  A lag the GPU chose may
    differ from the oracle's own only where the two lags' FP64 NCC values tie
    (within 1e-6); the oracle then refines from the GPU's lags."""
    steps = cfg2_fit["steps"]
    assert len(steps) > it  # from random labels the fit runs past iteration 2
    Z64 = cfg2_fit["Z"].astype(np.float64)
    prev_lab = cfg2_fit["init"] if it == 0 else steps[it - 1]["lab"]
    prev_C = np.zeros((K, M)) if it == 0 else steps[it - 1]["C"].astype(np.float64)
    shifts = None
    if it > 0:
        shifts = steps[it - 1]["shift"]
        ties = 0
        for j in range(K):
            rows = np.nonzero(prev_lab == j)[0]
            if rows.size == 0 or not np.any(prev_C[j]):
                continue
            _, _, own = O.assign(Z64[rows], prev_C[j:j + 1])
            for q in np.nonzero(own != shifts[rows])[0]:
                cc = O.ncc(Z64[rows[q]], prev_C[j])
                a, b = cc[own[q] + M - 1], cc[shifts[rows[q]] + M - 1]
                assert abs(a - b) <= SBD_TIE, (j, int(rows[q]), a, b)
                ties += 1
        print(f"iteration {it + 1}: {ties} alignment lags differ from the oracle's by a tie",
              flush=True)
    Co, counts, _ = O.refine(Z64, prev_lab, prev_C, seed=SEED, it=it, shifts=shifts)
    errs = centroid_errors(steps[it]["C"], Co)
    print(f"iteration {it + 1}: centroid rel-L2 max {max(errs):.3e}", flush=True)
    assert max(errs) <= CENT_RTOL, errs
    olab, _, _, odm = O.assign(Z64, steps[it]["C"].astype(np.float64), dmat=True)
    assert_labels_match(steps[it]["lab"], olab, odm, what=f"iteration {it + 1}")
