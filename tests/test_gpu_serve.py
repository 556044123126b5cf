"""Serving loop: every fit and query returns, in a mix submitted at once, the
result it gives alone; resident models equal the fits that publish them;
admission under an HBM budget; request errors stay with their request."""
import numpy as np
import pytest

import timesc as T

pytestmark = pytest.mark.gpu

FITS = [(300, 64, 3), (500, 128, 5), (200, 700, 2), (800, 96, 8)]


def _queries(rng, m, count):
    out = []
    for q in range(count):
        n = int(rng.integers(1, 300))
        X, _ = T.generate(int(rng.integers(1 << 30)), n, m, 3, 0.3)
        out.append(X)
    return out


def test_mix_equals_requests_alone(ctx):
    rng = np.random.default_rng(0)
    fits = [(T.generate(40 + i, n, m, k, 0.3)[0], k) for i, (n, m, k) in enumerate(FITS)]
    models = []
    for mm, kk in ((64, 4), (300, 12)):
        C, _ = T.generate(7 + mm, kk, mm, kk, 0.3)
        models.append(((C - C.mean(1, keepdims=True)) / C.std(1, keepdims=True)).astype(np.float32))
    queries = [(j, X) for j in range(2) for X in _queries(rng, models[j].shape[1], 12)]
    order = rng.permutation(len(fits) + len(queries))
    with T.Server(0, fit_workers=2, query_workers=2) as srv:
        mids = [srv.add_model(C) for C in models]
        handles = {}
        for idx in order:
            if idx < len(fits):
                X, k = fits[idx]
                handles[idx] = srv.submit_fit(X, k, seed=idx)
            else:
                j, X = queries[idx - len(fits)]
                handles[idx] = srv.submit_assign(mids[j], X)
        res = {idx: srv.wait(h) for idx, h in handles.items()}
    for i, (X, k) in enumerate(fits):
        d = ctx.to_device(X)
        try:
            ctx.znorm(d, ddof=0)
            alone = ctx.fit(d, k, seed=i)
        finally:
            d.free()
        assert res[i]["status"] == 0
        assert np.array_equal(res[i]["labels"], alone["labels"])
        assert np.array_equal(res[i]["centroids"], alone["centroids"])
    for q, (j, X) in enumerate(queries):
        d, dc = ctx.to_device(X), ctx.to_device(models[j])
        try:
            ctx.znorm(d, ddof=0)
            lab, dist, sh = ctx.assign(d, dc)
        finally:
            d.free()
            dc.free()
        r = res[len(fits) + q]
        assert np.array_equal(r["labels"], lab)
        assert np.array_equal(r["dist"], dist)
        assert np.array_equal(r["shift"], sh)
        assert 0 <= r["ms_queued"] <= r["ms_latency"]


def test_published_model_serves_the_fitted_centroids(ctx):
    X, _ = T.generate(3, 400, 80, 4, 0.3)
    Q, _ = T.generate(4, 150, 80, 4, 0.3)
    with T.Server(0, fit_workers=1, query_workers=1) as srv:
        f = srv.wait(srv.submit_fit(X, 4, publish=True, seed=1))
        assert f["model_out"] == 0
        q = srv.wait(srv.submit_assign(f["model_out"], Q))
    d, dc = ctx.to_device(Q), ctx.to_device(f["centroids"])
    try:
        ctx.znorm(d, ddof=0)
        lab, dist, _ = ctx.assign(d, dc)
    finally:
        d.free()
        dc.free()
    assert np.array_equal(q["labels"], lab) and np.array_equal(q["dist"], dist)


def test_admission_under_a_budget():
    X = [T.generate(60 + i, 1000, 128, 3, 0.3)[0] for i in range(4)]  # 0.52 MB each
    budget = 1_000_000  # room for one 0.524 MB request, not two
    with T.Server(0, fit_workers=3, query_workers=0, budget=budget) as srv:
        with pytest.raises(T.TimescError) as e:
            srv.submit_fit(T.generate(1, 4000, 128, 3, 0.3)[0], 3)  # 2 MB > budget
        assert e.value.code == T.ERR_BUDGET
        hs = [srv.submit_fit(x, 3, seed=1) for x in X]
        res = [srv.wait(h) for h in hs]
    assert all(r["status"] == 0 for r in res)
    # admitted one at a time: the runs (start, end), measured from nearly
    # simultaneous submissions, do not overlap
    spans = sorted((r["ms_queued"], r["ms_latency"]) for r in res)
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        assert s1 >= e0 - 1.0, spans


def test_request_errors_stay_with_their_request():
    Q, _ = T.generate(5, 20, 50, 2, 0.3)
    C, _ = T.generate(6, 2, 64, 2, 0.3)
    with T.Server(0, fit_workers=1, query_workers=1) as srv:
        mid = srv.add_model(C)
        bad_model = srv.submit_assign(mid + 5, Q)
        bad_len = srv.submit_assign(mid, Q)  # m = 50 against a model of m = 64
        ok = srv.submit_assign(mid, T.generate(7, 20, 64, 2, 0.3)[0])
        r1 = srv.wait(bad_model, raise_on_error=False)
        r2 = srv.wait(bad_len, raise_on_error=False)
        r3 = srv.wait(ok)
    assert r1["status"] == T.ERR_ARG and "unknown model" in r1["message"]
    assert r2["status"] == T.ERR_ARG and "differs" in r2["message"]
    assert r3["status"] == 0
