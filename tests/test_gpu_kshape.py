"""GPU parity of shape extraction and of the full k-Shape loop against the
FP64 oracle, through the C ABI."""
import numpy as np
import pytest

import oracle_py as O
import timesc as T
from parity import CENT_RTOL, assert_labels_match, centroid_errors, rel_l2

pytestmark = pytest.mark.gpu


def _znormed(ctx, seed, n, m, K, sigma=0.3, frac=None):
    X, y = T.generate(seed, n, m, K, sigma, frac)
    d = ctx.to_device(X)
    ctx.znorm(d, ddof=0)
    return d, d.numpy(), y


@pytest.mark.parametrize("m,n,k", [(128, 512, 4), (48, 300, 6), (512, 2000, 16), (300, 700, 5),
                                   (1000, 600, 3)])
@pytest.mark.parametrize("aligned", [False, True])
def test_extract_shape_matches_oracle(ctx, m, n, k, aligned):
    d, Z, y = _znormed(ctx, m + k, n, m, k)
    labels = T.init_labels(5, n, k) if not aligned else y.astype(np.int32)
    labels[labels == k - 1] = 0  # cluster k-1 empty -> zero centroid
    if aligned:
        C0 = O.zscore(Z[T.init_rows(9, n, k)].astype(np.float64), ddof=1).astype(np.float32)
    else:
        C0 = np.zeros((k, m), np.float32)
    dl, dc = ctx.to_device(labels), ctx.to_device(C0)
    try:
        sc = ctx.extract_shape(dl, d, None, dc, it=2)
        got = dc.numpy()
    finally:
        for a in (dl, dc, d):
            a.free()
    ref, counts, infos = O.refine(Z.astype(np.float64), labels, C0.astype(np.float64), seed=0, it=2)
    assert sc == int(np.sum(counts > 0))
    assert np.all(got[k - 1] == 0)
    errs = centroid_errors(got, ref)
    assert max(errs) <= CENT_RTOL, errs


@pytest.mark.parametrize("m", [64, 300])
def test_extract_shape_edge_clusters(ctx, m):
    """single-member cluster -> zscore_ddof1(member); constant members -> 0."""
    n = 10
    rng = np.random.default_rng(3)
    X = rng.standard_normal((n, m)).astype(np.float32)
    X[7:] = 1.5  # cluster 2: constant rows, z-normalized to zeros
    Z = O.zscore(X.astype(np.float64)).astype(np.float32)
    labels = np.array([0, 0, 0, 0, 0, 0, 1, 2, 2, 2], np.int32)
    d, dl = ctx.to_device(Z), ctx.to_device(labels)
    dc = ctx.to_device(np.zeros((3, m), np.float32))
    try:
        ctx.extract_shape(dl, d, None, dc)
        got = dc.numpy()
    finally:
        for a in (d, dl, dc):
            a.free()
    one = O.zscore(Z[6].astype(np.float64), ddof=1)
    assert rel_l2(got[1], one) <= CENT_RTOL / 10
    assert np.all(got[2] == 0)
    assert abs(got[0].mean()) <= 1e-5 and abs(got[0].std(ddof=1) - 1) <= 1e-5


@pytest.mark.parametrize("n,m,frac", [
    (20000, 256, [0.4, 0.3, 0.15, 0.1, 0.05]),  # many chunks of the 16-step minimum
    (30000, 1024, [0.5, 0.33, 0.17]),           # chunk bound searched, uneven chunk counts
    (6000, 2100, [0.6, 0.4]),                   # one (cluster, chunk) group wider than an XCD
])
def test_extract_shape_chunked_s_matches_oracle(ctx, n, m, frac):
    """Clusters large enough that S = Y^T Y is split into several K chunks
    per cluster, each summed from its own partials: the centroids match the
    oracle's; the last of k+1 centroids has no members and stays zero.  Both
    sides align every member by the oracle's FP64 shift, so that a member
    whose two best shifts tie within FP32 rounding (one in cluster 1 at
    m=256: 9.6e-8 apart) cannot move a centroid by a different alignment."""
    k = len(frac)
    d, Z, y = _znormed(ctx, 40 + k, n, m, k, frac=frac)
    labels = y.astype(np.int32)
    C0 = np.zeros((k + 1, m), np.float32)
    C0[:k] = O.zscore(Z[T.init_rows(11, n, k)].astype(np.float64), ddof=1)
    Z64, C64 = Z.astype(np.float64), C0.astype(np.float64)
    shifts = np.array([O.align_shift(Z64[i], C64[labels[i]]) for i in range(n)], np.int32)
    dl, dc, ds = ctx.to_device(labels), ctx.to_device(C0), ctx.to_device(shifts)
    try:
        sc = ctx.extract_shape(dl, d, ds, dc, it=3)
        got = dc.numpy()
    finally:
        for a in (dl, dc, ds, d):
            a.free()
    ref, counts, _ = O.refine(Z64, labels, C64, seed=0, it=3, shifts=shifts)
    assert sc == k and counts[k] == 0
    assert np.all(got[k] == 0)
    errs = centroid_errors(got, ref)
    assert max(errs) <= CENT_RTOL, errs


def test_stepwise_parity_cfg0(ctx):
    """Each refine and assign step from identical inputs (the GPU's previous
    state) matches the oracle: centroids to 1e-4 rel-L2, labels up to ties."""
    n, m, k = 512, 128, 4
    d, Z, y = _znormed(ctx, 0, n, m, k)
    Z64 = Z.astype(np.float64)
    labels = T.init_labels(0, n, k)
    C = np.zeros((k, m), np.float32)
    dl, dc = ctx.to_device(labels), ctx.to_device(C)
    try:
        for it in range(12):
            C_old = dc.numpy().astype(np.float64)
            lab_in = dl.numpy()
            ctx.extract_shape(dl, d, None, dc, it=it)
            Cg = dc.numpy()
            Co, _, _ = O.refine(Z64, lab_in, C_old, seed=0, it=it)
            errs = centroid_errors(Cg, Co)
            assert max(errs) <= CENT_RTOL, (it, errs)
            lab, dist, sh = ctx.assign(d, dc)
            olab, odist, osh, odm = O.assign(Z64, Cg.astype(np.float64), dmat=True)
            assert_labels_match(lab, olab, odm, what=f"iteration {it}")
            ctx.h2d(dl, lab)
    finally:
        for a in (d, dl, dc):
            a.free()


def test_fit_cfg0_trajectory(ctx):
    """Plain k-Shape, n=512, m=128, k=4, random-index init, 20 iterations:
    labels after every iteration equal the oracle's; final centroids within
    1e-4 rel-L2."""
    n, m, k, iters = 512, 128, 4, 20
    d, Z, y = _znormed(ctx, 0, n, m, k)
    init = T.init_labels(0, n, k)
    try:
        res = ctx.fit(d, k, max_iter=iters, init_labels=init, seed=0, labels_per_iter=True)
    finally:
        d.free()
    hist, cent = O.kshape(Z.astype(np.float64), k, init_labels=init, max_iter=iters, seed=0)
    got = res["labels_per_iter"]
    for t in range(min(len(got), len(hist))):
        if not np.array_equal(got[t], hist[t]):
            # only a near-tie may separate the trajectories
            C_prev = None
            raise AssertionError(f"trajectory diverges at iteration {t}: "
                                 f"{np.sum(got[t] != hist[t])} labels differ")
    assert len(got) == len(hist)
    errs = centroid_errors(res["centroids"], cent)
    assert max(errs) <= CENT_RTOL, errs


def test_fit_tsl_cfg1_small(ctx):
    """Centre initialisation (init_chose=1), long series with F=8192:
    n=137, m=2709, k=2, 60/40 classes."""
    n, m, k, iters = 137, 2709, 2, 8
    d, Z, y = _znormed(ctx, 1, n, m, k, frac=[0.6, 0.4])
    rows = T.init_rows(1, n, k)
    try:
        res = ctx.fit(d, k, max_iter=iters, init_mode=1, init_rows=rows, seed=1,
                      labels_per_iter=True)
    finally:
        d.free()
    hist, cent = O.kshape(Z.astype(np.float64), k, init_mode=1, init_rows=rows,
                          max_iter=iters, seed=1)
    got = res["labels_per_iter"]
    assert len(got) == len(hist)
    assert np.array_equal(got, hist)
    errs = centroid_errors(res["centroids"], cent)
    assert max(errs) <= CENT_RTOL, errs


@pytest.mark.parametrize("init_mode", [0, 1])
def test_fit_stops_on_collapse_like_the_gpu_loop(ctx, init_mode):
    """Two-class data fitted with k = 4 collapses to exactly k/2 non-empty
    clusters.  stop_on_collapse (the reference GPU loop's break,
    kshape.cuh:248-249, 454-455) stops right after that refinement with the
    labels of the assignment before it, as the oracle does; the
    reference-shaped tcx_kshape / tcx_kshape_tsl stop there too, return
    one-based labels, sum_class = k/2 and leave iter_num at the cap."""
    import ctypes

    n, m, k, seed = 60, 32, 4, 0  # seed 0: the reference-shaped entry points' default
    d, Z, _ = _znormed(ctx, 2, n, m, 2, sigma=0.1)
    kw = ({"init_labels": T.init_labels(seed, n, k)} if init_mode == 0
          else {"init_rows": T.init_rows(seed, n, k)})
    try:
        res = ctx.fit(d, k, max_iter=10, init_mode=init_mode, seed=seed, labels_per_iter=True,
                      stop_on_collapse=True, **kw)
        it, sc = ctypes.c_int(10), ctypes.c_int(-1)
        idx = np.empty(n, np.int32)
        fn = ctx.lib.tcx_kshape if init_mode == 0 else ctx.lib.tcx_kshape_tsl
        args = ([ctx.h, d.ptr] + ([None] if init_mode == 1 else []) +
                [idx.ctypes.data, n, m, k, ctypes.byref(it), ctypes.byref(sc), 0])
        T.check(fn(*args))
    finally:
        d.free()
    hist, _, col = O.kshape(Z.astype(np.float64), k, init_mode=init_mode, max_iter=10, seed=seed,
                            stop_on_collapse=True, return_collapsed=True, **kw)
    assert col and res["stats"]["collapsed"] == 1
    assert np.array_equal(res["labels_per_iter"], hist)
    assert np.array_equal(res["labels"], hist[-1])
    assert res["stats"]["sum_class"] == k // 2
    assert np.array_equal(idx, hist[-1] + 1) and sc.value == k // 2 and it.value == 10


def test_lanczos_wait_limit_falls_back_to_power_iteration(ctx, monkeypatch):
    """The multi-block Lanczos (m = 512, 8 clusters x 16 member blocks)
    with a negative wait limit set from the host: every member gives up at
    its first wait and exits, the clusters report "not accepted" and run
    the stepwise power iteration.  One refinement from the seed's labels:
    both fits' centroids are within the parity tolerance of the FP64
    oracle's refinement, and each fit's labels are the FP64 assignment of
    its own centroids up to SBD ties."""
    from parity import assert_labels_match

    n, m, k = 6000, 512, 8
    X, _ = T.generate(31, n, m, k, 0.3)

    def fit():
        d = ctx.to_device(X)
        try:
            ctx.znorm(d, ddof=0)
            r = ctx.fit(d, k, max_iter=1, seed=31)
            Z = d.numpy().astype(np.float64)
        finally:
            d.free()
        return r, Z

    ref, Z = fit()
    monkeypatch.setenv("TCX_LZ_COOP_TIMEOUT_US", "-1")
    got, _ = fit()
    print(f"power steps: Lanczos {ref['stats']['power_steps_max']}, forced "
          f"{got['stats']['power_steps_max']}", flush=True)
    # the stepwise iteration ran: up to min(m, 1000) steps, stopping at its
    # own tolerance, far more than the Lanczos steps
    assert ref["stats"]["power_steps_max"] < 64
    assert 64 < got["stats"]["power_steps_max"] <= m
    cent, _, _ = O.refine(Z, T.init_labels(31, n, k), np.zeros((k, m)), seed=31, it=0)
    for r, what in ((got, "forced fallback"), (ref, "Lanczos fit")):
        errs = [rel_l2(a, b) for a, b in zip(r["centroids"], cent) if np.any(b)]
        print(f"  {what}: centroids vs oracle max rel L2 {max(errs):.2e}", flush=True)
        assert max(errs) <= CENT_RTOL, (what, max(errs))
        olab, _, _, odm = O.assign(Z, r["centroids"].astype(np.float64), dmat=True)
        assert_labels_match(r["labels"], olab, odm, what=what)

