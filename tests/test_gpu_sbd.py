"""GPU parity of the SBD path (z-norm, NCC, assignment, alignment) against
the FP64 oracle, through the C ABI."""
import numpy as np
import pytest

import oracle_py as O
import timesc as T
from parity import SBD_TIE, assert_labels_match

pytestmark = pytest.mark.gpu

# one length per supported FFT size F = 1 << (2m-1).bit_length()
LENGTHS = [2, 3, 5, 9, 17, 33, 65, 128, 200, 512, 1000, 2048, 2709, 4096]


def _data(seed, n, m, k, sigma=0.3):
    X, y = T.generate(seed, n, m, k, sigma)
    return X, y


# m = 77 takes the three-pass kernel; multiples of 4 up to 4096 the one-pass
# register kernel (512, 1000, 2048, 4096: 2, 4, 8 and 16 float4 per lane)
@pytest.mark.parametrize("m", [77, 512, 1000, 2048, 4096])
@pytest.mark.parametrize("ddof", [0, 1])
def test_znorm_matches_zscore(ctx, ddof, m):
    X, _ = _data(1, 300, m, 3)
    X[5, :] = 2.0  # constant row -> zeros (nan_to_num)
    X[6, :] = 0.0
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=ddof)
        got = d.numpy()
    finally:
        d.free()
    ref = O.zscore(X.astype(np.float64), ddof=ddof)
    assert np.abs(got - ref).max() <= 2e-6
    assert np.all(got[5] == 0) and np.all(got[6] == 0)


@pytest.mark.parametrize("m", LENGTHS)
def test_ncc_matches_oracle(ctx, m):
    rng = np.random.default_rng(m)
    xr, yr = 5, 3
    X = rng.standard_normal((xr, m)).astype(np.float32)
    Y = rng.standard_normal((yr, m)).astype(np.float32)
    Y[2] = 0.0  # zero norm -> NCC 0 everywhere
    dx, dy = ctx.to_device(X), ctx.to_device(Y)
    try:
        got = ctx.ncc(dx, dy)
    finally:
        dx.free()
        dy.free()
    assert got.shape == (xr, yr, 2 * m - 1)
    for i in range(xr):
        for j in range(yr):
            ref = O.ncc(X[i].astype(np.float64), Y[j].astype(np.float64))
            assert np.abs(got[i, j] - ref).max() <= 2e-6, (i, j)
    assert np.all(got[:, 2, :] == 0)


@pytest.mark.parametrize("m,n,k", [(2, 31, 3), (17, 101, 4), (128, 512, 4), (512, 1001, 16),
                                   (1000, 77, 5), (2709, 40, 2), (4096, 9, 3)])
def test_assign_matches_oracle(ctx, m, n, k):
    X, _ = _data(m + n, n, m, k)
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        Z = d.numpy()
        rows = T.init_rows(3, n, k)
        C = Z[rows].copy()
        if k > 2:
            C[k - 1] = 0.0  # zero centroid: distance exactly 1
        dc = ctx.to_device(C)
        lab, dist, sh, dm = ctx.assign(d, dc, dmat=True)
        dc.free()
    finally:
        d.free()
    olab, odist, osh, odm = O.assign(Z.astype(np.float64), C.astype(np.float64), dmat=True)
    err = np.abs(dm - odm).max()
    assert err <= 2e-6, f"max SBD error {err:.3e}"
    assert_labels_match(lab, olab, odm)
    same = lab == olab
    assert np.abs(dist[same] - odist[same]).max() <= 2e-6
    if k > 2:
        assert np.all(dm[:, k - 1] == 1.0)
    # shifts: equal unless the oracle's NCC has a near-tie between the lags
    for i in np.nonzero(same & (sh != osh))[0]:
        cc = O.ncc(Z[i].astype(np.float64), C[lab[i]].astype(np.float64))
        assert abs(cc[sh[i] + m - 1] - cc[osh[i] + m - 1]) <= SBD_TIE, (i, sh[i], osh[i])


@pytest.mark.parametrize("m,n,k", [(2, 31, 3), (2, 40, 90), (17, 101, 4), (60, 333, 40),
                                   (128, 512, 4), (512, 1001, 16), (512, 600, 130),
                                   (1000, 77, 5), (2709, 40, 7), (4096, 9, 3)])
def test_pruned_assign_matches_oracle(ctx, m, n, k):
    """The assignment without a distance matrix skips centroids by their
    spectral bound; labels, distances and shifts stay those of the full
    search (k=90 at m=2 exceeds the pruned kernel's LDS and takes the full
    search)."""
    X, _ = _data(2 * m + n, n, m, min(k, 8))
    X[3] = 0.0  # an all-zero series: distance 1 to everything, label 0
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        Z = d.numpy()
        rng = np.random.default_rng(m * k)
        rows = rng.choice(n, size=k, replace=k > n)
        C = Z[rows].copy()
        if k > 2:
            C[1] = 0.0          # zero centroid: distance exactly 1
            C[k - 1] = C[k - 2]  # duplicate centroids: the smaller index wins
        dc = ctx.to_device(C)
        lab, dist, sh = ctx.assign(d, dc)
        dc.free()
    finally:
        d.free()
    olab, odist, osh, odm = O.assign(Z.astype(np.float64), C.astype(np.float64), dmat=True)
    assert_labels_match(lab, olab, odm)
    assert lab[3] == 0 and dist[3] == 1.0 and sh[3] == 1 - m
    if k > 2:
        assert not np.any(lab == k - 1)
    same = lab == olab
    assert np.abs(dist[same] - odist[same]).max() <= 2e-6
    for i in np.nonzero(same & (sh != osh))[0]:
        cc = O.ncc(Z[i].astype(np.float64), C[lab[i]].astype(np.float64))
        assert abs(cc[sh[i] + m - 1] - cc[osh[i] + m - 1]) <= SBD_TIE, (i, sh[i], osh[i])


# F = 256 (16 lanes), 1024 (one wave: m = 260 leaves 505 non-lag positions
# over rows 4..11, m = 512 one), 2048, 4096 and 8192 (two, four and eight
# waves per group); k up to 64 keeps a one-wave group's bounds in
# registers, k = 80 in LDS
@pytest.mark.parametrize("m,n,k", [(96, 3001, 12), (2048, 257, 6), (260, 2001, 16),
                                   (512, 3001, 16), (300, 1501, 64), (400, 801, 80),
                                   (700, 1201, 12), (1030, 601, 9), (3000, 301, 5)])
def test_pruned_assign_equals_full_search(ctx, m, n, k):
    """Pruned and full assignment of the same inputs agree bit for bit on
    labels and distances (both pair rows 2p, 2p+1 in one transform)."""
    X, _ = _data(5 * m + k, n, m, k)
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        Z = d.numpy()
        C = Z[T.init_rows(7, n, k)].copy()
        dc = ctx.to_device(C)
        lab, dist, sh = ctx.assign(d, dc)
        flab, fdist, fsh, _ = ctx.assign(d, dc, dmat=True)
        dc.free()
    finally:
        d.free()
    assert np.array_equal(lab, flab)
    assert np.array_equal(dist, fdist)
    assert np.array_equal(sh, fsh)


@pytest.mark.parametrize("m", [260, 512, 700, 2048])
def test_pruned_assign_equals_full_search_on_ties(ctx, m):
    """Series and centroids of +-1 values: many exactly equal correlations
    (across lags and across centroids), so the first-key and first-label
    rules decide; pruned and full search agree bit for bit."""
    n, k = 1201, 12
    rng = np.random.default_rng(m)
    C = rng.choice([-1.0, 1.0], size=(k, m)).astype(np.float32)
    X = C[rng.integers(0, k, n)] * rng.choice([-1.0, 1.0], size=(n, 1)).astype(np.float32)
    flip = rng.random((n, m)) < 0.2
    X = np.where(flip, -X, X).astype(np.float32)
    X[::7, : m // 2] = 1.0  # constant halves: flat correlation plateaus
    d = ctx.to_device(X)
    dc = ctx.to_device(C)
    try:
        lab, dist, sh = ctx.assign(d, dc)
        flab, fdist, fsh, _ = ctx.assign(d, dc, dmat=True)
    finally:
        dc.free()
        d.free()
    assert np.array_equal(lab, flab)
    assert np.array_equal(dist, fdist)
    assert np.array_equal(sh, fsh)


# (m, k) over the bound product's forms: row steps of 16 columns per
# batch 9 (m = 100, F = 256), 17 (m = 200), 33 (m = 512), 33 + 32 (m = 700,
# F = 2048); 1, 2 and 4 centroid tiles (k <= 16, <= 32, > 32)
@pytest.mark.parametrize("m,k", [(100, 5), (100, 24), (100, 40), (200, 12), (512, 8), (700, 20)])
def test_fit_assignment_equals_dense_search(ctx, m, k):
    """The assignment inside a fit (series half spectra once per fit, the
    bound rows in binary16 rounded up, the bound product per iteration)
    labels every series as the dense search over the same centroids, up to
    ties."""
    n = 600
    X, _ = _data(11 * m + k, n, m, k)
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        ctx.fit_begin(d, k, max_iter=1 << 30, seed=3)
        for _ in range(3):
            ctx.fit_step()
        res = ctx.fit_end(n, k, m)
        dc = ctx.to_device(res["centroids"])
        try:
            flab, _, _, fdm = ctx.assign(d, dc, dmat=True)
        finally:
            dc.free()
    finally:
        d.free()
    assert_labels_match(res["labels"], flab, fdm, what=f"fit assignment m={m} k={k}")


@pytest.mark.parametrize("convention", [T.CONV_CPU, T.CONV_KSHAPE])
@pytest.mark.parametrize("m", [16, 128, 700])
def test_sbd_align_matches_oracle(ctx, convention, m):
    X, _ = _data(m, 40, m, 2)
    Z = O.zscore(X.astype(np.float64)).astype(np.float32)
    c = Z[0].copy()
    members = Z[1:].copy()
    dm_, dc = ctx.to_device(members), ctx.to_device(c[None, :])
    try:
        ctx.sbd_align(dm_, dc, convention)
        got = dm_.numpy()
        # all-zero centre: members unchanged (test_timesC.py:128)
        dz = ctx.to_device(np.zeros((1, m), np.float32))
        dm2 = ctx.to_device(members)
        ctx.sbd_align(dm2, dz, convention)
        unchanged = dm2.numpy()
        dz.free()
        dm2.free()
    finally:
        dm_.free()
        dc.free()
    assert np.array_equal(unchanged, members)
    for i in range(members.shape[0]):
        s = O.align_shift(members[i].astype(np.float64), c.astype(np.float64), convention)
        ref = O.roll_zeropad(members[i].astype(np.float64), s)
        assert np.array_equal(got[i], ref.astype(np.float32)), i


def test_assign_known_shift(ctx):
    """member = roll_zeropad(c, d): the CPU-path shift is +d."""
    m = 256
    X, _ = _data(11, 1, m, 1)
    c = O.zscore(X[0].astype(np.float64)).astype(np.float32)
    d = 13
    mem = O.roll_zeropad(c.astype(np.float64), d).astype(np.float32)
    dx, dc = ctx.to_device(np.stack([mem, c])), ctx.to_device(c[None, :])
    try:
        lab, dist, sh = ctx.assign(dx, dc)
    finally:
        dx.free()
        dc.free()
    assert sh[0] == d and sh[1] == 0
    assert abs(dist[1]) <= 1e-6
