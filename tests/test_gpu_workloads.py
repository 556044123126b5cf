"""The batched and served workloads at their own sizes, through the C ABI:

  - 256 uneven independent fits (n ~ U{100..5000}, m = round(exp U[ln 64,
    ln 2048]), k ~ U{2..32}, seeds 1000 + j, at most 10 iterations), run as
    one batch: every job's labels, centroids and iteration count equal the
    same fit run alone on one context;
  - a Poisson request stream (400 requests at 100/s: 30% fits of the sizes
    above, 70% assignment queries of 1..1024 series against 8 resident
    models with m <= 1024 and k <= 64) replayed at its arrival times on a
    server: every result equals the request run alone.
The workloads are the benchmark's own (bench.batch_workload and
bench.serve_workload)."""
import numpy as np
import pytest

import bench
import timesc as T

pytestmark = pytest.mark.gpu


def _alone_fit(ctx, X, k, seed):
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        return ctx.fit(d, k, max_iter=10, seed=seed)
    finally:
        d.free()


@pytest.fixture(scope="module")
def batch_jobs():
    print("\ngenerating the 256-job workload", flush=True)
    return bench.batch_workload(T, 256)


def test_batch_of_256_uneven_jobs_equals_each_job_alone(ctx, batch_jobs):
    jobs = batch_jobs
    ds = [ctx.to_device(X) for X, _, _ in jobs]
    try:
        with T.Batch(0, workers=8) as b:
            res = b.run([{"d": d, "k": k, "seed": seed, "znorm_ddof": 0, "max_iter": 10}
                         for d, (_, k, seed) in zip(ds, jobs)], order=T.ORDER_LONGEST_FIRST)
    finally:
        for d in ds:
            d.free()
    print("batch done; running every job alone", flush=True)
    for j, ((X, k, seed), r) in enumerate(zip(jobs, res)):
        assert r["status"] == 0, (j, r["message"])
        a = _alone_fit(ctx, X, k, seed)
        assert np.array_equal(r["labels"], a["labels"]), j
        assert np.array_equal(r["centroids"], a["centroids"]), j
        assert r["stats"]["iterations"] == a["stats"]["iterations"], j


def test_poisson_serving_stream_equals_each_request_alone(ctx):
    models, reqs = bench.serve_workload(T, 400, 100.0)
    assert max(C.shape[0] for C in models) <= 64 and max(C.shape[1] for C in models) <= 1024
    print(f"\nreplaying {len(reqs)} requests at 100/s", flush=True)
    import time

    with T.Server(0, fit_workers=8, query_workers=4) as srv:
        mids = [srv.add_model(C) for C in models]
        t0 = time.perf_counter()
        hs = []
        for t, kind, X, a in reqs:
            delay = t0 + t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            hs.append(srv.submit_fit(X, a) if kind == "fit" else srv.submit_assign(mids[a], X))
        res = [srv.wait(h) for h in hs]
    print("stream served; running every request alone", flush=True)
    dmodels = [ctx.to_device(C) for C in models]
    try:
        for i, ((t, kind, X, a), r) in enumerate(zip(reqs, res)):
            assert r["status"] == 0, (i, r["message"])
            if kind == "fit":
                alone = _alone_fit(ctx, X, a, 0)
                assert np.array_equal(r["labels"], alone["labels"]), i
                assert np.array_equal(r["centroids"], alone["centroids"]), i
            else:
                d = ctx.to_device(X)
                try:
                    ctx.znorm(d, ddof=0)
                    lab, dist, sh = ctx.assign(d, dmodels[a])
                finally:
                    d.free()
                assert np.array_equal(r["labels"], lab), i
                assert np.array_equal(r["dist"], dist), i
                assert np.array_equal(r["shift"], sh), i
    finally:
        for d in dmodels:
            d.free()
