"""The autotuner (tcx_autotune) through the C ABI: its launch shapes never
change a result.  Fits and grouped batches under forced extreme entries of
the table (the smallest grid cap, every align + pack row count the LDS
allows) equal the same runs with the built-in shapes, bit for bit; a tuning
run returns valid choices and leaves a table that reproduces the default
results."""
import numpy as np
import pytest

import timesc as T

pytestmark = pytest.mark.gpu


def _fit(ctx, X, k, seed=5):
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        return ctx.fit(d, k, max_iter=6, seed=seed, labels_per_iter=True)
    finally:
        d.free()


def _grouped(ctx, data):
    ds = [ctx.to_device(X) for X, _ in data]
    try:
        with T.Batch(0, workers=1) as b:
            return b.run([dict(d=d, k=k, seed=3, znorm_ddof=0, max_iter=6, labels_per_iter=True)
                          for d, (_, k) in zip(ds, data)], grouped=True)
    finally:
        for d in ds:
            d.free()


def _same(a, b):
    assert np.array_equal(a["labels_per_iter"], b["labels_per_iter"])
    assert np.array_equal(a["centroids"], b["centroids"])
    assert a["stats"]["sbd_evals"] == b["stats"]["sbd_evals"]


@pytest.mark.parametrize("n,m,k", [(3000, 128, 6), (2000, 300, 12), (1500, 1000, 20)])
def test_forced_shapes_equal_builtin(ctx, n, m, k):
    X, _ = T.generate(77, n, m, k, 0.3)
    data = [(X, k), (T.generate(78, 700, m, 3, 0.3)[0], 3)]
    T.tune_reset()
    try:
        ref = _fit(ctx, X, k)
        ref_g = _grouped(ctx, data)
        F = 1 << (2 * m - 1).bit_length()
        m32 = (m + 31) // 32 * 32
        rows = [r for r in (16, 32, 64) if r * (m32 + 4) * 4 <= 72 * 1024 and m32 <= 1024]
        for cap in (64, 100000):
            for r in rows or [0]:
                T.tune_reset()
                for nn, kk in ((n, k), (700, 3)):  # the fit and both grouped jobs
                    T.tune_set(T.TUNE_PRUNE_CAP, F, (kk + 15) // 16, nn.bit_length() - 1, cap)
                if r:
                    T.tune_set(T.TUNE_PACK_ROWS, m32, 0, 0, r)
                print(f"n={n} m={m} k={k}: cap {cap}, rows {r}", flush=True)
                _same(_fit(ctx, X, k), ref)
                for a, b in zip(_grouped(ctx, data), ref_g):
                    _same(a, b)
    finally:
        T.tune_reset()


def test_autotune_choices_reproduce_default(ctx, tmp_path):
    n, m, k = 6000, 256, 8
    X, _ = T.generate(79, n, m, k, 0.3)
    T.tune_reset()
    try:
        ref = _fit(ctx, X, k)
        d = ctx.to_device(X)
        try:
            ctx.znorm(d, ddof=0)
            rep = ctx.autotune(d, k, reps=1)
        finally:
            d.free()
        print("autotune:", rep, flush=True)
        assert rep["fft"] == 1 << (2 * m - 1).bit_length()
        assert rep["prune_cap"] in (0, 1024, 2048, 4096, 8192, 16384, 32768)
        assert rep["pack_rows"] in (0, 16, 32, 64)
        assert 0 < rep["ms_tuned"] <= rep["ms_default"]
        nb = n.bit_length() - 1
        assert T.tune_get(T.TUNE_PRUNE_CAP, rep["fft"], 1, nb) == rep["prune_cap"]
        assert T.tune_get(T.TUNE_PACK_ROWS, m) == rep["pack_rows"]
        _same(_fit(ctx, X, k), ref)
        path = tmp_path / "tune.txt"
        T.tune_save(path)
        T.tune_reset()
        T.tune_load(path)
        assert T.tune_get(T.TUNE_PRUNE_CAP, rep["fft"], 1, nb) == rep["prune_cap"]
    finally:
        T.tune_reset()
