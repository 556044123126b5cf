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


def _fit_need(ctx, n, m, k):
    """Bytes a fit request needs: staging (series, labels, shifts,
    distances) plus the fit's workspace bound."""
    return n * m * 4 + n * 12 + ctx.fit_workspace_bound(n, m, k)


def test_admission_under_a_budget(ctx):
    X = [T.generate(60 + i, 1000, 128, 3, 0.3)[0] for i in range(4)]
    need = _fit_need(ctx, 1000, 128, 3)
    budget = need + need // 2  # room for one such fit, not two
    with T.Server(0, fit_workers=3, query_workers=0, budget=budget) as srv:
        with pytest.raises(T.TimescError) as e:
            srv.submit_fit(T.generate(1, 8000, 128, 3, 0.3)[0], 3)  # needs > budget
        assert e.value.code == T.ERR_BUDGET
        hs = [srv.submit_fit(x, 3, seed=1) for x in X]
        res = [srv.wait(h) for h in hs]
    assert all(r["status"] == 0 for r in res)
    # admitted one at a time: the runs (start, end), measured from nearly
    # simultaneous submissions, do not overlap
    spans = sorted((r["ms_queued"], r["ms_latency"]) for r in res)
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        assert s1 >= e0 - 1.0, spans


def test_fit_workspace_bound_holds(ctx):
    """The bound the admission reserves is never below what a fit grows."""
    for i, (n, m, k) in enumerate([(1000, 128, 3), (3000, 700, 12), (257, 2, 2), (5000, 2048, 32),
                                   (600, 1500, 4), (100000, 512, 16)]):
        X, _ = T.generate(70 + i, n, m, k, 0.3)
        with T.Context(0) as c:
            d = c.to_device(X)
            c.znorm(d, ddof=0)
            before = c.workspace_bytes()
            c.fit(d, k, max_iter=4, seed=i)
            grown = c.workspace_bytes() - before
            bound = c.fit_workspace_bound(n, m, k)
            d.free()
        assert 0 < grown <= bound, (n, m, k, grown, bound)


def test_budget_below_one_fit_refuses_it_and_models_count(ctx):
    """A fit whose workspace cannot fit the budget is refused at submission;
    resident models reduce what is left, and a query that no longer fits is
    refused too; the server then stops cleanly."""
    n, m, k = 2000, 256, 4
    need = _fit_need(ctx, n, m, k)
    X, _ = T.generate(80, n, m, k, 0.3)
    with T.Server(0, fit_workers=1, query_workers=1, budget=need - 1) as srv:
        with pytest.raises(T.TimescError) as e:
            srv.submit_fit(X, k)
        assert e.value.code == T.ERR_BUDGET
    C, _ = T.generate(81, 64, 1024, 8, 0.3)  # a 64-centroid model of m = 1024
    Q, _ = T.generate(82, 300, 1024, 8, 0.3)
    with T.Server(0, fit_workers=1, query_workers=1, budget=2_200_000) as srv:
        mid = srv.add_model(C)  # 1.85 MB resident (centroids, spectra, twiddles)
        with pytest.raises(T.TimescError) as e:
            srv.submit_assign(mid, Q)  # 1.25 MB of staging no longer fits beside it
        assert e.value.code == T.ERR_BUDGET
        ok = srv.wait(srv.submit_assign(mid, Q[:50]))
        assert ok["status"] == 0


def test_destroy_fails_queued_requests():
    """Requests still queued when the server stops fail with an error; the
    running one finishes; destroy returns."""
    fits = [T.generate(90 + i, 3000, 300, 8, 0.3)[0] for i in range(6)]
    srv = T.Server(0, fit_workers=1, query_workers=0)
    hs = [srv.submit_fit(X, 8, max_iter=10) for X in fits]
    reqs = [h["req"] for h in hs]
    srv.close()
    done = [r.done for r in reqs]
    status = [r.status for r in reqs]
    assert all(done)
    assert status[0] == 0 or status[0] == T.ERR_ARG
    assert any(s == T.ERR_ARG for s in status[1:])
    failed = [r.message.decode() for r in reqs if r.status == T.ERR_ARG]
    assert all("stopped before the request ran" in msg for msg in failed)


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
