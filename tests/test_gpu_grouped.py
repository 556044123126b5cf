"""The grouped batch executor (tcx_batch_run_grouped: one launch per phase
over all jobs in progress), through the C ABI:

  - a job's results do not depend on the other jobs: every job of a mixed
    batch equals the same job run alone through the executor, bit for bit;
  - the assignment phase replayed as a HIP graph equals the same launches
    enqueued directly, every iteration;
  - parity: the labels after every iteration equal the FP64 oracle's and the
    final centroids are within 1e-4 relative L2, over shapes that take
    FFT lengths 16..4096, the S and the Gram form of the power stage, every
    align + pack form and both bound-product forms, on seeds whose oracle
    trajectories hold no SBD or alignment-lag near-tie; and the first
    iteration of large clusters (every power-stage form) against the
    oracle's refinement and assignment;
  - the collapse stop, and jobs the executor leaves to the worker pool
    (centre initialisation, eigenvector centroids) in the same call;
  - a batch object reused for different job sets gives each set's
    fresh-batch results (the buffers kept between calls carry nothing);
  - cohorts (batches of 32 jobs or more split over up to
    TCX_GROUPED_COHORTS contexts, with a second stream created or borrowed
    from an idle worker): every layout gives the one-cohort results bit for
    bit;
  - the 256-job uneven workload (two cohorts): every job equals itself alone
    through the executor, and its final labels and centroids equal
    tcx_kshape_fit's (to rounding; differing labels only at SBD ties)."""
import numpy as np
import pytest

import bench
import oracle_py as O
import timesc as T
from parity import CENT_RTOL, rel_l2

pytestmark = pytest.mark.gpu

# (n, m, k), seed: F = 16 .. 4096; clusters shorter than m (in padded
# members, up to 1024) take the Gram form (one wave per cluster up to order
# 256, one block above), the others the S form (one wave for m <= 256, the
# two-launch form above: (1200, 300, 2)); m > 1024 takes the align + pack
# from HBM;
# k > 16 the wider bound product.  On these seeds the oracle's trajectory
# has no near-tie: in every
# iteration the best and second-best SBD of every series, and the best and
# second-best alignment lag of every member, differ by more than 2e-6
# (checked on the CPU when the seeds were chosen), so FP32 must reproduce the
# FP64 labels and centroids.  (m = 2 is left to the independence test:
# z-normalized rows of length 2 are +-(1, -1), every assignment is a tie.)
CASES = [((300, 30, 3), 710), ((500, 128, 5), 711), ((257, 5, 2), 712), ((400, 260, 17), 913),
         ((400, 700, 4), 1214), ((350, 900, 6), 1115), ((150, 1500, 3), 813),
         ((2000, 96, 16), 1417), ((64, 33, 5), 718), ((200, 2048, 2), 1413),
         ((1200, 300, 2), 1818)]
SHAPES = [c[0] for c in CASES]


def _specs(ctx, shapes, seeds, **kw):
    ds, specs, Xs = [], [], []
    for (n, m, k), seed in zip(shapes, seeds):
        X, _ = T.generate(seed, n, m, k, 0.3)
        d = ctx.to_device(X)
        ds.append(d)
        Xs.append(X)
        specs.append(dict(d=d, k=k, seed=seed, znorm_ddof=0, **{"max_iter": 10, **kw}))
    return ds, specs, Xs


def _run(ctx, shapes, seeds, grouped=True, workers=2, **kw):
    ds, specs, Xs = _specs(ctx, shapes, seeds, **kw)
    try:
        with T.Batch(0, workers=workers) as b:
            res = b.run(specs, grouped=grouped)
    finally:
        for d in ds:
            d.free()
    return res, Xs


def test_grouped_jobs_equal_themselves_alone(ctx):
    shapes = SHAPES + [(257, 2, 2), (800, 260, 17), (1200, 2048, 2)]
    seeds = [700 + i for i in range(len(shapes))]
    res, _ = _run(ctx, shapes, seeds)
    for i, shape in enumerate(shapes):
        alone, _ = _run(ctx, [shape], [seeds[i]])
        a, r = alone[0], res[i]
        assert r["status"] == 0 and a["status"] == 0, (r["message"], a["message"])
        assert np.array_equal(r["labels"], a["labels"]), shape
        assert np.array_equal(r["centroids"], a["centroids"]), shape
        assert r["stats"]["iterations"] == a["stats"]["iterations"], shape
        assert r["stats"]["sbd_evals"] == a["stats"]["sbd_evals"], shape


def test_grouped_assignment_graph_equals_direct_launches(ctx, monkeypatch):
    """The assignment phase replayed as a HIP graph (the default) gives the
    same labels after every iteration and the same centroids as the same
    launches enqueued directly (TCX_GROUPED_GRAPH=0): several replays per
    job, with the refinement's eager work between them."""
    shapes = [(300, 30, 3), (500, 128, 5), (400, 700, 4), (2000, 96, 16), (1200, 300, 2)]
    seeds = [31, 32, 33, 34, 35]
    graph, _ = _run(ctx, shapes, seeds, labels_per_iter=True)
    monkeypatch.setenv("TCX_GROUPED_GRAPH", "0")
    direct, _ = _run(ctx, shapes, seeds, labels_per_iter=True)
    for shape, g, d in zip(shapes, graph, direct):
        assert g["status"] == 0 and d["status"] == 0, (g["message"], d["message"])
        assert g["stats"]["iterations"] >= 3, shape
        assert np.array_equal(g["labels_per_iter"], d["labels_per_iter"]), shape
        assert np.array_equal(g["centroids"], d["centroids"]), shape
        assert g["stats"]["sbd_evals"] == d["stats"]["sbd_evals"], shape


def test_grouped_trajectories_match_oracle(ctx):
    seeds = [c[1] for c in CASES]
    res, Xs = _run(ctx, SHAPES, seeds, labels_per_iter=True)
    for (n, m, k), seed, r, X in zip(SHAPES, seeds, res, Xs):
        assert r["status"] == 0, r["message"]
        Z = O.zscore(X.astype(np.float64))
        hist, cent = O.kshape(Z, k, init_labels=T.init_labels(seed, n, k), max_iter=10,
                              seed=seed)
        got = r["labels_per_iter"]
        assert got.shape == hist.shape and np.array_equal(got, hist), (n, m, k)
        errs = [rel_l2(a, b) for a, b in zip(r["centroids"], cent) if np.any(b)]
        assert max(errs) <= CENT_RTOL, ((n, m, k), errs)


# First iteration (no alignment: the initial centroids are zero) of shapes
# that take each power-stage form with large clusters: fat clusters with
# m > 256 (the two-launch form, m <= 768 and above), thin clusters whose
# Gram order exceeds 256 (one block per cluster), and the one-wave form
# (Gram and S).  Free of alignment-lag ties, so any seed does.
# (2400, 1800, 2): clusters of about 1200 members, fewer than m but above
# the Gram block form's limit of 1024: the S form
FIRST_ITER = [((1200, 300, 2), 1801), ((1700, 780, 2), 1802), ((2500, 900, 2), 1803),
              ((1600, 1200, 2), 1804), ((700, 900, 2), 1805), ((600, 128, 4), 1806),
              ((900, 700, 6), 1807), ((2400, 1800, 2), 1808)]


def test_grouped_first_iteration_matches_oracle(ctx):
    shapes = [c[0] for c in FIRST_ITER]
    seeds = [c[1] for c in FIRST_ITER]
    res, Xs = _run(ctx, shapes, seeds, max_iter=1)
    from parity import assert_labels_match

    for (n, m, k), seed, r, X in zip(shapes, seeds, res, Xs):
        assert r["status"] == 0, r["message"]
        assert r["stats"]["iterations"] == 1
        Z = O.zscore(X.astype(np.float64))
        cent, _, _ = O.refine(Z, T.init_labels(seed, n, k), np.zeros((k, m)), seed=seed, it=0)
        errs = [rel_l2(a, b) for a, b in zip(r["centroids"], cent) if np.any(b)]
        assert max(errs) <= CENT_RTOL, ((n, m, k), errs)
        olab, _, _, odm = O.assign(Z, r["centroids"].astype(np.float64), dmat=True)
        assert_labels_match(r["labels"], olab, odm, what=f"{(n, m, k)}")


def test_grouped_collapse_stop_and_pool_jobs_in_one_call(ctx):
    """A two-class job with k = 4 stops on collapse like the single-fit
    engine; a centre-initialized job and an eigenvector-centroid job run on
    the worker pool in the same call."""
    X, _ = T.generate(2, 60, 32, 2, 0.1)
    Y, _ = T.generate(3, 400, 96, 4, 0.3)
    ds = [ctx.to_device(X), ctx.to_device(Y), ctx.to_device(Y)]
    specs = [dict(d=ds[0], k=4, seed=0, znorm_ddof=0, stop_on_collapse=True),
             dict(d=ds[1], k=4, seed=1, znorm_ddof=0, init_mode=1),
             dict(d=ds[2], k=4, seed=1, znorm_ddof=0, cal_select=1)]
    try:
        with T.Batch(0, workers=2) as b:
            res = b.run(specs, grouped=True)
    finally:
        for d in ds:
            d.free()
    assert all(r["status"] == 0 for r in res), [r["message"] for r in res]
    Z = O.zscore(X.astype(np.float64))
    hist, _, col = O.kshape(Z, 4, init_labels=T.init_labels(0, 60, 4), max_iter=10, seed=0,
                            stop_on_collapse=True, return_collapsed=True)
    assert col and res[0]["stats"]["collapsed"] == 1
    assert np.array_equal(res[0]["labels"], hist[-1])
    for r, kw in ((res[1], dict(init_mode=1)), (res[2], dict(cal_select=1))):
        d = ctx.to_device(Y)
        try:
            ctx.znorm(d, ddof=0)
            alone = ctx.fit(d, 4, max_iter=10, seed=1, **kw)
        finally:
            d.free()
        assert np.array_equal(r["labels"], alone["labels"])
        assert np.array_equal(r["centroids"], alone["centroids"])


@pytest.mark.parametrize("workers,cohorts", [(2, None), (3, "2"), (4, None), (8, "3")])
def test_grouped_cohorts_equal_one_cohort(ctx, monkeypatch, workers, cohorts):
    """40 jobs: two cohorts with a created second stream (2 and 3 workers),
    two with borrowed ones (4 workers), three with borrowed ones (8)."""
    shapes = [(120 + 37 * i, (24, 64, 130, 300)[i % 4], 2 + i % 5) for i in range(40)]
    seeds = [3000 + i for i in range(40)]
    monkeypatch.setenv("TCX_GROUPED_COHORTS", "1")
    ref, _ = _run(ctx, shapes, seeds, workers=1, labels_per_iter=True)
    if cohorts is None:
        monkeypatch.delenv("TCX_GROUPED_COHORTS")
    else:
        monkeypatch.setenv("TCX_GROUPED_COHORTS", cohorts)
    print(f"{workers} workers, cohorts {cohorts or 'default'}", flush=True)
    res, _ = _run(ctx, shapes, seeds, workers=workers, labels_per_iter=True)
    for i, (a, b) in enumerate(zip(res, ref)):
        assert a["status"] == 0 and b["status"] == 0, (i, a["message"], b["message"])
        assert np.array_equal(a["labels_per_iter"], b["labels_per_iter"]), i
        assert np.array_equal(a["centroids"], b["centroids"]), i
        assert a["stats"]["iterations"] == b["stats"]["iterations"], i


def test_grouped_reused_batch_equals_fresh_batches(ctx):
    """One batch object runs job set A (40 jobs: two cohorts), then a smaller
    set B of other shapes, then A again: the executor's buffers, kept in
    the worker contexts between calls, carry nothing from one run into the
    next; every run equals the same set on a fresh batch, bit for bit."""
    sa = [(150 + 41 * i, (40, 96, 200, 320)[i % 4], 2 + i % 6) for i in range(40)]
    sb = [(900 + 13 * i, (700, 64, 1300)[i % 3], 3 + i % 4) for i in range(7)]
    seeds_a = [5000 + i for i in range(40)]
    seeds_b = [6000 + i for i in range(7)]
    da, specs_a, _ = _specs(ctx, sa, seeds_a, labels_per_iter=True)
    db, specs_b, _ = _specs(ctx, sb, seeds_b, labels_per_iter=True)
    Xa = [d.numpy() for d in da]
    Xb = [d.numpy() for d in db]

    def reload():
        for d, X in zip(da + db, Xa + Xb):
            ctx.h2d(d, X)

    try:
        fresh = []
        for specs in (specs_a, specs_b):
            reload()
            with T.Batch(0, workers=4) as b:
                fresh.append(b.run(specs, grouped=True))
        with T.Batch(0, workers=4) as b:
            runs = []
            for specs in (specs_a, specs_b, specs_a):
                reload()
                runs.append(b.run(specs, grouped=True))
    finally:
        for d in da + db:
            d.free()
    for got, want in zip(runs, (fresh[0], fresh[1], fresh[0])):
        for i, (a, r) in enumerate(zip(got, want)):
            assert a["status"] == 0 and r["status"] == 0, (i, a["message"], r["message"])
            assert np.array_equal(a["labels_per_iter"], r["labels_per_iter"]), i
            assert np.array_equal(a["centroids"], r["centroids"]), i


def test_grouped_256_uneven_jobs(ctx):
    """Every job: its final labels are the FP64 oracle's assignment against
    its own final centroids (up to SBD ties); 32 sampled jobs equal
    themselves run alone through the executor; at least 85% end with the
    single-fit engine's labels.  The rest differ through FP32 ties met on
    the way, in SBD or in an alignment lag: the executor evaluates the power
    iterate of a thin cluster through its Gram matrix, the single-fit engine
    through S, so the two round differently, and ten iterations carry a
    flipped near-tie forward."""
    from parity import assert_labels_match

    jobs = bench.batch_workload(T, 256)
    ds = [ctx.to_device(X) for X, _, _ in jobs]
    specs = [dict(d=d, k=k, seed=seed, znorm_ddof=0, max_iter=10)
             for d, (_, k, seed) in zip(ds, jobs)]
    try:
        with T.Batch(0, workers=4) as b:  # two cohorts
            res = b.run(specs, grouped=True)
    finally:
        for d in ds:
            d.free()
    print("grouped batch done; checking every job", flush=True)
    same_labels = 0
    rng = np.random.default_rng(0)
    alone_check = set(rng.choice(len(jobs), 32, replace=False).tolist())
    for i, ((X, k, seed), r) in enumerate(zip(jobs, res)):
        assert r["status"] == 0, (i, r["message"])
        d = ctx.to_device(X)
        try:
            ctx.znorm(d, ddof=0)
            Z = d.numpy().astype(np.float64)
            single = ctx.fit(d, k, max_iter=10, seed=seed)
        finally:
            d.free()
        olab, _, _, odm = O.assign(Z, r["centroids"].astype(np.float64), dmat=True)
        assert_labels_match(r["labels"], olab, odm, what=f"job {i}")
        same_labels += int(np.array_equal(r["labels"], single["labels"]))
        if i in alone_check:
            d = ctx.to_device(X)
            try:
                with T.Batch(0, workers=1) as b:
                    a = b.run([dict(d=d, k=k, seed=seed, znorm_ddof=0, max_iter=10)],
                              grouped=True)[0]
            finally:
                d.free()
            assert np.array_equal(r["labels"], a["labels"]), i
            assert np.array_equal(r["centroids"], a["centroids"]), i
    print(f"{same_labels} of {len(jobs)} jobs end with the single-fit engine's labels")
    assert same_labels >= 0.85 * len(jobs)
