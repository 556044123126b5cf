"""The grouped batch executor (tcx_batch_run_grouped: one launch per phase
over all jobs in progress), through the C ABI:

  - a job's results do not depend on the other jobs: every job of a mixed
    batch equals the same job run alone through the executor, bit for bit;
  - parity: the labels after every iteration equal the FP64 oracle's and the
    final centroids are within 1e-4 relative L2, over shapes that take every
    kernel form (FFT lengths 64..4096, the fused and the two-launch power
    stage, every align + pack form, both bound-product forms);
  - the collapse stop, and jobs the executor leaves to the worker pool
    (centre initialisation, eigenvector centroids) in the same call;
  - the 256-job uneven workload: every job equals itself alone through the
    executor, and its final labels and centroids equal tcx_kshape_fit's
    (to rounding; differing labels only at SBD ties)."""
import numpy as np
import pytest

import bench
import oracle_py as O
import timesc as T
from parity import CENT_RTOL, rel_l2

pytestmark = pytest.mark.gpu

# (n, m, k): F = 64 .. 4096; m > 768 takes the two-launch power stage,
# m > 1024 the align + pack from HBM; k > 16 the wider bound product
SHAPES = [(300, 30, 3), (500, 128, 5), (257, 2, 2), (900, 300, 17), (400, 700, 4),
          (350, 900, 6), (300, 1500, 3), (2000, 96, 16), (64, 33, 5), (1200, 2048, 2)]


def _specs(ctx, shapes, seed0, **kw):
    ds, specs, Xs = [], [], []
    for i, (n, m, k) in enumerate(shapes):
        X, _ = T.generate(seed0 + i, n, m, k, 0.3)
        d = ctx.to_device(X)
        ds.append(d)
        Xs.append(X)
        specs.append(dict(d=d, k=k, seed=seed0 + i, znorm_ddof=0, max_iter=10, **kw))
    return ds, specs, Xs


def _run(ctx, shapes, seed0, grouped=True, workers=2, **kw):
    ds, specs, Xs = _specs(ctx, shapes, seed0, **kw)
    try:
        with T.Batch(0, workers=workers) as b:
            res = b.run(specs, grouped=grouped)
    finally:
        for d in ds:
            d.free()
    return res, Xs


def test_grouped_jobs_equal_themselves_alone(ctx):
    res, _ = _run(ctx, SHAPES, 700)
    for i, shape in enumerate(SHAPES):
        alone, _ = _run(ctx, [shape], 700 + i)
        a, r = alone[0], res[i]
        assert r["status"] == 0 and a["status"] == 0, (r["message"], a["message"])
        assert np.array_equal(r["labels"], a["labels"]), shape
        assert np.array_equal(r["centroids"], a["centroids"]), shape
        assert r["stats"]["iterations"] == a["stats"]["iterations"], shape
        assert r["stats"]["sbd_evals"] == a["stats"]["sbd_evals"], shape


def test_grouped_trajectories_match_oracle(ctx):
    res, Xs = _run(ctx, SHAPES, 710, labels_per_iter=True)
    for i, ((n, m, k), r, X) in enumerate(zip(SHAPES, res, Xs)):
        assert r["status"] == 0, r["message"]
        Z = O.zscore(X.astype(np.float64))
        hist, cent = O.kshape(Z, k, init_labels=T.init_labels(710 + i, n, k), max_iter=10,
                              seed=710 + i)
        got = r["labels_per_iter"]
        assert got.shape == hist.shape and np.array_equal(got, hist), (n, m, k)
        errs = [rel_l2(a, b) for a, b in zip(r["centroids"], cent) if np.any(b)]
        assert max(errs) <= CENT_RTOL, ((n, m, k), errs)


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


def test_grouped_256_uneven_jobs(ctx):
    jobs = bench.batch_workload(T, 256)
    ds = [ctx.to_device(X) for X, _, _ in jobs]
    specs = [dict(d=d, k=k, seed=seed, znorm_ddof=0, max_iter=10)
             for d, (_, k, seed) in zip(ds, jobs)]
    try:
        with T.Batch(0, workers=1) as b:
            res = b.run(specs, grouped=True)
    finally:
        for d in ds:
            d.free()
    print("grouped batch done; checking every job alone", flush=True)
    same_labels = 0
    rng = np.random.default_rng(0)
    alone_check = set(rng.choice(len(jobs), 32, replace=False).tolist())
    for i, ((X, k, seed), r) in enumerate(zip(jobs, res)):
        assert r["status"] == 0, (i, r["message"])
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
        d = ctx.to_device(X)
        try:
            ctx.znorm(d, ddof=0)
            single = ctx.fit(d, k, max_iter=10, seed=seed)
        finally:
            d.free()
        if np.array_equal(r["labels"], single["labels"]):
            same_labels += 1
            errs = [rel_l2(a, b) for a, b in zip(r["centroids"], single["centroids"]) if np.any(b)]
            assert max(errs) <= CENT_RTOL, (i, errs)
    print(f"{same_labels} of {len(jobs)} jobs end with the single-fit engine's labels")
    assert same_labels >= 0.9 * len(jobs)
