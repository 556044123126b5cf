"""tcx_assign_noalloc (static device scratch, no allocation) against
tcx_assign, through the C ABI: two calls in a row, the second with a smaller
input, labels, distances, lags and the distance matrix bit for bit; sizes
beyond the static scratch are refused."""
import numpy as np
import pytest

import timesc as T

pytestmark = pytest.mark.gpu


def _inputs(seed, n, m, k):
    X, _ = T.generate(seed, n, m, k, 0.4)
    C, _ = T.generate(seed + 1, k, m, k, 0.4)
    return X, C


def _run(ctx, cases, dmat):
    """All cases through tcx_assign_noalloc back to back, then through
    tcx_assign; the results must be identical."""
    devs = []
    try:
        for X, C in cases:
            d, dc = ctx.to_device(X), ctx.to_device(C)
            devs += [d, dc]
            ctx.znorm(d, ddof=0)
            ctx.znorm(dc, ddof=1)
        pairs = list(zip(devs[0::2], devs[1::2]))
        got = [ctx.assign(d, dc, dmat=dmat, noalloc=True) for d, dc in pairs]
        want = [ctx.assign(d, dc, dmat=dmat) for d, dc in pairs]
    finally:
        for d in devs:
            d.free()
    for i, (g, w) in enumerate(zip(got, want)):
        for a, b in zip(g, w):
            assert np.array_equal(a, b), (i, dmat)


@pytest.mark.parametrize("dmat", [False, True])
def test_noalloc_assign_equals_assign_twice_second_smaller(ctx, dmat):
    # F = 1024 then a smaller n, m and k (F = 256)
    _run(ctx, [_inputs(11, 5000, 512, 16), _inputs(12, 700, 100, 5)], dmat)


def test_noalloc_assign_k_beyond_pruned_lds_and_long_series(ctx):
    # F = 8192 (k * F = 65536), then k = 1000 at F = 128
    _run(ctx, [_inputs(21, 300, 2709, 8), _inputs(22, 200, 40, 1000)], False)


def test_noalloc_assign_refuses_sizes_beyond_the_scratch(ctx):
    X = np.zeros((4, 1024), np.float32)
    C = np.zeros((200, 1024), np.float32)  # k * F = 200 * 2048 > 262144
    d, dc = ctx.to_device(X), ctx.to_device(C)
    try:
        with pytest.raises(T.TimescError) as e:
            ctx.assign(d, dc, noalloc=True)
    finally:
        d.free()
        dc.free()
    assert e.value.code == T.ERR_ARG and "static scratch" in str(e.value)
