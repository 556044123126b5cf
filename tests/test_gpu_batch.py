"""Batched fits (worker pool of contexts): every job's result equals the
same fit run alone on one context, for every job order; a failing job
reports its own error without disturbing the others."""
import numpy as np
import pytest

import timesc as T

pytestmark = pytest.mark.gpu

SHAPES = [(300, 64, 3), (257, 2, 2), (1000, 128, 8), (150, 700, 4), (64, 33, 5), (2000, 96, 16),
          (500, 1100, 2), (120, 17, 3), (800, 256, 12), (90, 512, 6)]


def _inputs():
    out = []
    for i, (n, m, k) in enumerate(SHAPES):
        X, _ = T.generate(500 + i, n, m, k, 0.3)
        out.append((X, k, 500 + i))
    return out


@pytest.fixture(scope="module")
def alone(ctx):
    res = []
    for X, k, seed in _inputs():
        d = ctx.to_device(X)
        try:
            ctx.znorm(d, ddof=0)
            res.append(ctx.fit(d, k, max_iter=10, seed=seed))
        finally:
            d.free()
    return res


@pytest.mark.parametrize("order", [T.ORDER_GIVEN, T.ORDER_SHORTEST_FIRST, T.ORDER_LONGEST_FIRST])
def test_batch_equals_single_fits(ctx, alone, order):
    inputs = _inputs()
    ds = [ctx.to_device(X) for X, _, _ in inputs]
    try:
        with T.Batch(0, workers=4) as b:
            res = b.run([{"d": d, "k": k, "seed": seed, "znorm_ddof": 0, "max_iter": 10}
                         for d, (_, k, seed) in zip(ds, inputs)], order=order)
    finally:
        for d in ds:
            d.free()
    workers = set()
    for r, a in zip(res, alone):
        assert r["status"] == 0, r["message"]
        assert np.array_equal(r["labels"], a["labels"])
        assert np.array_equal(r["centroids"], a["centroids"])
        assert r["stats"]["iterations"] == a["stats"]["iterations"]
        assert 0 <= r["ms_queued"] <= r["ms_latency"] and r["ms_run"] > 0
        workers.add(r["worker"])
    assert workers <= {0, 1, 2, 3} and len(workers) > 1


def test_batch_reports_a_failing_job(ctx):
    X, _ = T.generate(9, 200, 40, 3, 0.3)
    d1, d2 = ctx.to_device(X), ctx.to_device(X)
    try:
        with T.Batch(0, workers=2) as b:
            res = b.run([{"d": d1, "k": 3, "znorm_ddof": 0},
                         {"d": d2, "k": 300, "init_mode": 1, "znorm_ddof": 0}],
                        raise_on_error=False)
            with pytest.raises(T.TimescError) as e:
                b.run([{"d": d2, "k": 300, "init_mode": 1}])
    finally:
        d1.free()
        d2.free()
    assert res[0]["status"] == 0 and res[1]["status"] == T.ERR_ARG
    assert "k > n" in res[1]["message"]
    assert e.value.code == T.ERR_ARG and "job 0" in str(e.value)
