"""The precision of S = Y^T Y through the C ABI (tcx_fit_opts.s_precision):

  - TCX_S_AUTO runs split-BF16 until a refinement's BF16 centroids are
    within s_budget of its split ones, then BF16.  With the default budget
    (1e-4) these shapes never switch (BF16 S moves their centroids by
    1.6e-4 to 1e-3) and the fit is the split fit bit for bit; with a budget
    of 5e-3 they switch at iteration 2: until then the labels are the split
    fit's bit for bit (the check adds work, changes nothing), and the last
    refinement's centroids stay within twice the budget of the FP64
    oracle's refinement of the fit's own previous labels;
  - TCX_S_BF16 reports BF16 from the first iteration."""
import numpy as np
import pytest

import oracle_py as O
import timesc as T
from parity import rel_l2

pytestmark = pytest.mark.gpu

SHAPES = [(6000, 256, 8, 21), (4000, 512, 6, 22), (3000, 128, 12, 23)]


def _fit(ctx, X, k, seed, **kw):
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        return ctx.fit(d, k, max_iter=10, seed=seed, labels_per_iter=True, **kw)
    finally:
        d.free()


@pytest.mark.parametrize("n,m,k,seed", SHAPES)
def test_auto_precision_default_budget_keeps_split(ctx, n, m, k, seed):
    X, _ = T.generate(seed, n, m, k, 0.3)
    split = _fit(ctx, X, k, seed)
    auto = _fit(ctx, X, k, seed, s_precision=T.S_AUTO)
    err = auto["stats"]["s_check_err"]
    print(f"n={n} m={m} k={k}: BF16 - split centroids {err:.2e}", flush=True)
    assert split["stats"]["s_bf16_from"] == -1
    assert auto["stats"]["s_bf16_from"] == -1 and err > 1e-4
    assert np.array_equal(auto["labels_per_iter"], split["labels_per_iter"])
    assert np.array_equal(auto["centroids"], split["centroids"])


@pytest.mark.parametrize("n,m,k,seed", SHAPES)
def test_auto_precision_switches_within_budget(ctx, n, m, k, seed):
    budget = 5e-3
    X, _ = T.generate(seed, n, m, k, 0.3)
    split = _fit(ctx, X, k, seed)
    auto = _fit(ctx, X, k, seed, s_precision=T.S_AUTO, s_budget=budget)
    s = auto["stats"]["s_bf16_from"]
    print(f"n={n} m={m} k={k}: BF16 from iteration {s}", flush=True)
    assert s == 2
    assert np.array_equal(auto["labels_per_iter"][:s], split["labels_per_iter"][:s])
    # the last refinement against the FP64 oracle's, from the fit's own
    # labels before it (the labels of the previous assignment)
    hist = auto["labels_per_iter"]
    it = auto["stats"]["iterations"]
    assert it > s  # the last refinement ran in BF16
    Z = O.zscore(X.astype(np.float64))
    prev = hist[it - 2]
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        ctx.fit_begin(d, k, max_iter=10, seed=seed, s_precision=T.S_AUTO, s_budget=budget)
        for _ in range(it - 1):
            ctx.fit_step()
        lab, dist, shift, cent_prev = ctx.fit_snapshot(n, k, m)
        ctx.fit_step()
        _, _, _, cent = ctx.fit_snapshot(n, k, m)
        ctx.fit_end(n, k, m)
    finally:
        d.free()
    assert np.array_equal(lab, prev)
    ref, _, _ = O.refine(Z, lab, cent_prev.astype(np.float64), seed=seed, it=it - 1,
                         shifts=shift)
    errs = [rel_l2(a, b) for a, b in zip(cent, ref) if np.any(b)]
    print(f"  last refinement vs oracle: max rel L2 {max(errs):.2e}", flush=True)
    assert max(errs) <= 2 * budget


def test_bf16_reports_its_precision(ctx):
    X, _ = T.generate(24, 1500, 200, 4, 0.3)
    r = _fit(ctx, X, 4, 24, s_precision=T.S_BF16)
    assert r["stats"]["s_bf16_from"] == 0
