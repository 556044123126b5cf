"""Pins the FP64 CPU oracle (oracle/oracle.c) against independent checks:
NumPy's FFT and eigh, O(m^2) brute-force correlation, closed forms, known
answers, and an independent NumPy restatement of the reference's `_kshape`
loop (Code/test/test_timesC.py:213-265).  CPU only."""
import math

import numpy as np
import pytest

import oracle_py as O
import timesc as T


# ----------------------------------------------------------------- helpers
def fft_size(m):
    return 1 << (2 * m - 1).bit_length()


def direct_ncc(x, y):
    """O(m^2) cross-correlation cc[l] = sum_u x[u+l] y[u], l = -(m-1)..m-1,
    divided by |x||y| (0 when a norm is 0)."""
    m = len(x)
    cc = np.array([sum(x[u + l] * y[u] for u in range(m) if 0 <= u + l < m)
                   for l in range(-(m - 1), m)], dtype=np.float64)
    den = np.linalg.norm(x) * np.linalg.norm(y)
    return cc / den if den != 0 else np.zeros_like(cc)


def np_zscore(a, ddof=0):
    a = np.asarray(a, np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        mu = a.mean(axis=-1, keepdims=True)
        sd = a.std(axis=-1, ddof=ddof, keepdims=True)
        return np.nan_to_num((a - mu) / sd)


def np_roll_zeropad(a, s):
    n = a.size
    out = np.zeros_like(a)
    if abs(s) >= n:
        return out
    if s >= 0:
        out[s:] = a[:n - s]
    else:
        out[:n + s] = a[-s:]
    return out


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("F", [2 ** p for p in range(1, 15)])
def test_fft_matches_numpy(F):
    rng = np.random.default_rng(F)
    x = rng.standard_normal(F) + 1j * rng.standard_normal(F)
    scale = np.abs(x).sum()
    assert np.abs(O.fft(x) - np.fft.fft(x)).max() <= 1e-14 * scale
    assert np.abs(O.fft(x, inverse=True) / F - np.fft.ifft(x)).max() <= 1e-14 * scale / F


def test_fft_impulse_and_sinusoid():
    F = 64
    x = np.zeros(F, complex)
    x[0] = 1
    assert np.allclose(O.fft(x), np.ones(F), atol=1e-15)
    t = np.arange(F)
    X = O.fft(np.exp(2j * np.pi * 5 * t / F))
    ref = np.zeros(F)
    ref[5] = F
    assert np.abs(X - ref).max() <= 1e-12


@pytest.mark.parametrize("m,F", [(1, 2), (2, 4), (3, 8), (128, 256), (512, 1024), (2709, 8192),
                                 (4096, 8192)])
def test_fft_size_rule(m, F):
    # 1 << (2m-1).bit_length() (test_timesC.py:110)
    assert O.fft_size(m) == F == fft_size(m)


# ---------------------------------------------------------------- zscore
@pytest.mark.parametrize("ddof", [0, 1])
def test_zscore_closed_form(ddof):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((20, 37)) * 3 + 1
    got = O.zscore(X, ddof=ddof)
    assert np.abs(got - np_zscore(X, ddof)).max() <= 1e-13
    assert np.abs(got.mean(axis=1)).max() <= 1e-13
    assert np.abs(got.std(axis=1, ddof=ddof) - 1).max() <= 1e-13


def test_zscore_constant_rows_are_zero():
    X = np.array([[2.0] * 16, [0.0] * 16, [-0.5] * 16])
    assert np.all(O.zscore(X, 0) == 0) and np.all(O.zscore(X, 1) == 0)


# ----------------------------------------------------------- roll_zeropad
def test_roll_zeropad_closed_form():
    a = np.arange(1.0, 12.0)
    for s in range(-14, 15):
        assert np.array_equal(O.roll_zeropad(a, s), np_roll_zeropad(a, s)), s
    assert np.all(O.roll_zeropad(a, a.size) == 0) and np.all(O.roll_zeropad(a, -a.size) == 0)


# -------------------------------------------------------------------- NCC
@pytest.mark.parametrize("m", [1, 2, 3, 7, 16, 33, 64])
def test_ncc_matches_direct(m):
    rng = np.random.default_rng(m)
    x, y = rng.standard_normal(m), rng.standard_normal(m)
    assert np.abs(O.ncc(x, y) - direct_ncc(x, y)).max() <= 1e-13


def test_ncc_invariants():
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(50), rng.standard_normal(50)
    c = O.ncc(x, x)
    assert abs(c[49] - 1) <= 1e-14 and np.all(np.abs(O.ncc(x, y)) <= 1 + 1e-14)
    assert np.all(O.ncc(x, np.zeros(50)) == 0)


@pytest.mark.parametrize("d", [-7, -1, 1, 3, 12])
def test_align_shift_known_answer(d):
    """member = roll_zeropad(c, d): the CPU path shifts by +d
    (test_timesC.py:129-132), the k-Shape/GPU path by -d (SBD.cuh:61,78)."""
    rng = np.random.default_rng(abs(d))
    t = np.arange(80)
    c = np.sin(2 * np.pi * 3 * t / 80) + np.exp(-(t - 30) ** 2 / 20) + 0.05 * rng.standard_normal(80)
    member = np_roll_zeropad(c, d)
    assert O.align_shift(member, c, O.CONV_CPU) == d
    assert O.align_shift(member, c, O.CONV_KSHAPE) == -d


# ------------------------------------------------------------- assignment
def test_assign_matches_brute_force():
    rng = np.random.default_rng(9)
    n, m, k = 25, 19, 4
    X = np_zscore(rng.standard_normal((n, m)))
    C = np_zscore(rng.standard_normal((k, m)))
    C[2] = 0.0  # empty-cluster centroid: distance exactly 1
    lab, dist, sh, dm = O.assign(X, C, dmat=True)
    for i in range(n):
        d = np.array([1 - direct_ncc(X[i], C[j]).max() for j in range(k)])
        assert np.abs(dm[i] - d).max() <= 1e-13
        assert lab[i] == int(np.argmin(d))
        assert sh[i] == int(np.argmax(direct_ncc(X[i], C[lab[i]]))) + 1 - m
    assert np.all(dm[:, 2] == 1.0)


def test_assign_ties_take_first_index():
    rng = np.random.default_rng(2)
    X = np_zscore(rng.standard_normal((5, 16)))
    C = np.stack([X[1], X[1], X[3]])  # duplicate centroids 0 and 1
    lab, _, _ = O.assign(X, C)
    assert lab[1] == 0 and lab[3] == 2


# ------------------------------------------------------- shape extraction
def closed_form_centroid(A):
    """zscore_ddof1(signfix_{a0}(top eigenvector of PSP)), test_timesC.py:141-204
    with eigh in place of the converged power iteration."""
    m = A.shape[1]
    y = np_zscore(A, 1)
    S = y.T @ y
    P = np.eye(m) - np.full((m, m), 1.0 / m)
    w, V = np.linalg.eigh(P @ S @ P)
    c = V[:, -1]
    if math.sqrt(((A[0] - c) ** 2).sum()) >= math.sqrt(((A[0] + c) ** 2).sum()):
        c = -c
    return np_zscore(c, 1)


def test_extract_shape_matches_eigh():
    rng = np.random.default_rng(4)
    t = np.arange(60)
    proto = np.sin(2 * np.pi * 2 * t / 60) + 0.5 * np.cos(2 * np.pi * 5 * t / 60)
    A = np.stack([a * proto + 0.2 * rng.standard_normal(60) for a in rng.uniform(0.5, 2, 40)])
    got, info = O.extract_shape(A, O.power_v0(0, 0, 0, 60))
    ref = closed_form_centroid(A)
    assert np.abs(got - ref).max() <= 1e-9
    assert abs(got.mean()) <= 1e-12 and abs(got.std(ddof=1) - 1) <= 1e-12
    assert A[0] @ got > 0 and info["fallback"] == 0


def test_extract_shape_eigenvector_path_is_the_eigh_branch():
    """cal_select = 1: no power iteration, the top eigenvector of PSP with the
    sign fix (test_timesC.py:184-192), even where the power iterate after
    min(m, 1000) steps is still far from it."""
    rng = np.random.default_rng(8)
    t = np.arange(48)
    p1, p2 = np.sin(2 * np.pi * 3 * t / 48), np.cos(2 * np.pi * 4 * t / 48)
    A = np.stack([rng.normal(1, 0.3) * p1 + rng.normal(0, 0.98) * p2 + 0.05 * rng.standard_normal(48)
                  for _ in range(30)])
    got, info = O.extract_shape(A, O.power_v0(0, 0, 0, 48), cal_select=1)
    assert np.abs(got - closed_form_centroid(A)).max() <= 1e-9
    assert info["fallback"] == 1 and info["steps"] == 0


def test_extract_shape_one_member_and_degenerate():
    rng = np.random.default_rng(7)
    a = rng.standard_normal((1, 32))
    got, _ = O.extract_shape(a, O.power_v0(0, 0, 0, 32))
    assert np.abs(got - np_zscore(a[0], 1)).max() <= 1e-10
    # constant members: NumPy's power iteration divides 0 by 0, the NaN
    # centroid becomes zeros through zscore's nan_to_num
    got, info = O.extract_shape(np.full((3, 32), 1.5), O.power_v0(0, 0, 0, 32))
    assert np.all(got == 0)
    got, info = O.extract_shape(np.zeros((0, 32)), O.power_v0(0, 0, 0, 32))
    assert np.all(got == 0) and info["empty"] == 1


def test_top_eig_matches_numpy():
    rng = np.random.default_rng(11)
    for m in (2, 5, 40, 130):
        B = rng.standard_normal((m, m))
        M = B @ B.T
        lam, vec = O.top_eig(M)
        w, V = np.linalg.eigh(M)
        assert abs(lam - w[-1]) <= 1e-10 * abs(w[-1])
        assert min(np.abs(vec - V[:, -1]).max(), np.abs(vec + V[:, -1]).max()) <= 1e-8


def test_power_v0_matches_engine():
    for it, cl, m in [(0, 0, 5), (3, 7, 64), (11, 2, 513)]:
        assert np.array_equal(O.power_v0(123, it, cl, m), T.power_v0(123, it, cl, m))
        v = O.power_v0(123, it, cl, m)
        assert np.all((v >= 0) & (v < 1))


# ---------------------------------------------- the loop, restated in NumPy
def np_refine(X, idx, C, seed, it):
    """One refinement of _kshape (test_timesC.py:228-255) in NumPy: align each
    cluster to its old centroid (all-zero rule), power iteration of exactly
    min(m, 1000) steps from power_v0, sign fix, zscore.  Returns the new
    centroids and the cluster sizes."""
    n, m = X.shape
    k = C.shape[0]
    F = fft_size(m)
    C = C.copy()
    counts = np.array([np.sum(idx == j) for j in range(k)])
    for j in range(k):
        A = X[idx == j].copy()
        if len(A) == 0:
            C[j] = 0
            continue
        if np.any(C[j] != 0):  # :128-132, all-zero rule (see test below)
            for q in range(len(A)):
                cc = np.real(np.fft.ifft(np.fft.fft(A[q], F) * np.conj(np.fft.fft(C[j], F))))
                cc = np.concatenate((cc[-(m - 1):], cc[:m]))
                den = np.linalg.norm(A[q]) * np.linalg.norm(C[j])
                cc = cc / den if den != 0 else np.zeros_like(cc)
                A[q] = np_roll_zeropad(A[q], int(np.argmax(cc)) + 1 - m)
        y = np_zscore(A, 1)
        S = y.T @ y
        P = np.eye(m) - np.full((m, m), 1.0 / m)
        M = P @ S @ P
        v = O.power_v0(seed, it, j, m)
        v = v / np.linalg.norm(v)
        for _ in range(min(m, 1000)):
            v = M @ v
            v = v / np.linalg.norm(v)
        lam = (M @ v) @ v / (v @ v)
        if math.sqrt(((A[0] - v) ** 2).sum()) >= math.sqrt(((A[0] + v) ** 2).sum()):
            v = -v
        assert lam > 0
        C[j] = np_zscore(v, 1)
    return C, counts


def np_assign(X, C):
    """dist = 1 - max_lag NCC, label = first argmin (test_timesC.py:258-260)."""
    n, m = X.shape
    F = fft_size(m)
    Xf = np.fft.fft(X, F)
    D = np.empty((n, C.shape[0]))
    for j in range(C.shape[0]):
        cc = np.real(np.fft.ifft(Xf * np.conj(np.fft.fft(C[j], F))))
        cc = np.concatenate((cc[:, -(m - 1):], cc[:, :m]), axis=1)
        den = np.linalg.norm(X, axis=1) * np.linalg.norm(C[j])
        with np.errstate(invalid="ignore", divide="ignore"):
            nc = np.where(den[:, None] != 0, cc / den[:, None], 0.0)
        D[:, j] = 1 - nc.max(axis=1)
    return D.argmin(axis=1).astype(np.int32)


def collapsed(counts, k):
    """sum_class == 1 || sum_class <= k/2 (kshape.cuh:248, 454)."""
    s = int(np.sum(counts > 0))
    return s == 1 or s <= k // 2


def np_kshape(X, k, init_labels, max_iter, seed, stop_on_collapse=False):
    """Independent restatement of _kshape (test_timesC.py:213-265) with the
    engine's injected initial labels and power start vectors; with
    stop_on_collapse, the GPU loop's break after extract_shape
    (kshape.cuh:248-249)."""
    n, m = X.shape
    idx = init_labels.copy()
    C = np.zeros((k, m))
    hist = []
    for it in range(max_iter):
        old = idx.copy()
        C, counts = np_refine(X, idx, C, seed, it)
        if stop_on_collapse and collapsed(counts, k):
            break
        idx = np_assign(X, C)
        hist.append(idx.copy())
        if np.array_equal(old, idx):
            break
    return np.array(hist).reshape(-1, n), C


def np_kshape_tsl(X, k, rows, max_iter, seed, stop_on_collapse=False):
    """Independent restatement of k_shape_tsl's loop (kshape.cuh:376-461):
    centres = the chosen rows (z-normalized, as X is), then per iteration
    assign, stop when the labels equal the previous iteration's (from the
    second iteration on), refine, and the collapse break (:454-455)."""
    idx = None
    C = X[rows].copy()
    hist = []
    for it in range(max_iter):
        new = np_assign(X, C)
        hist.append(new.copy())
        if it > 0 and np.array_equal(idx, new):
            break
        idx = new
        C, counts = np_refine(X, idx, C, seed, it)
        if stop_on_collapse and collapsed(counts, k):
            break
    return np.array(hist), C


def test_kshape_matches_numpy_restatement():
    n, m, k, iters = 60, 32, 3, 6
    Xf, _ = T.generate(21, n, m, k, 0.3)
    X = O.zscore(Xf.astype(np.float64))
    init = T.init_labels(21, n, k)
    hist, cent = O.kshape(X, k, init_labels=init, max_iter=iters, seed=5, exact_steps=True)
    ref_hist, ref_cent = np_kshape(X, k, init, iters, 5)
    assert hist.shape == ref_hist.shape and np.array_equal(hist, ref_hist)
    assert np.abs(cent - ref_cent).max() <= 1e-8


@pytest.mark.parametrize("seed,n,m,k", [(31, 80, 40, 3), (32, 120, 24, 4), (33, 64, 70, 2)])
def test_kshape_tsl_matches_numpy_restatement(seed, n, m, k):
    """init_mode 1 (k_shape_tsl, kshape.cuh:376-461; no CPU-path counterpart):
    the oracle's loop equals a NumPy restatement built from brute-force
    assignment and eigh-checked refinement, iteration by iteration."""
    Xf, _ = T.generate(seed, n, m, k, 0.4)
    X = O.zscore(Xf.astype(np.float64))
    rows = T.init_rows(seed, n, k)
    hist, cent = O.kshape(X, k, init_mode=1, init_rows=rows, max_iter=8, seed=seed,
                          exact_steps=True)
    ref_hist, ref_cent = np_kshape_tsl(X, k, rows, 8, seed)
    assert hist.shape == ref_hist.shape and np.array_equal(hist, ref_hist)
    assert np.abs(cent - ref_cent).max() <= 1e-8


def test_kshape_tsl_first_assignment_is_brute_force_against_the_rows():
    """The first labels of init_mode 1 are the brute-force SBD argmin against
    the chosen rows themselves (O(m^2) correlation, no FFT)."""
    n, m, k = 40, 20, 3
    Xf, _ = T.generate(41, n, m, k, 0.5)
    X = O.zscore(Xf.astype(np.float64))
    rows = T.init_rows(41, n, k)
    hist, _ = O.kshape(X, k, init_mode=1, init_rows=rows, max_iter=1)
    want = [int(np.argmin([1 - direct_ncc(X[i], X[r]).max() for r in rows])) for i in range(n)]
    assert np.array_equal(hist[0], want)


def test_kshape_tsl_loop_order_known_answer():
    """Noise-free, shift-free, separated classes with one centre row per
    class: k_shape_tsl assigns the true labels, refines, assigns them again
    and stops -- two iterations (assign first), where the label-initialized
    loop started from the true labels stops after one (refine first)."""
    n, m, k = 30, 40, 3
    t = np.arange(m)
    protos = [np.sin(2 * np.pi * f * t / m) for f in (1, 3, 7)]
    y = np.repeat(np.arange(k), n // k).astype(np.int32)
    rng = np.random.default_rng(0)
    X = O.zscore(np.stack([rng.uniform(0.5, 2) * protos[c] + rng.normal() for c in y]))
    rows = np.array([0, 10, 20], np.int32)
    hist, _ = O.kshape(X, k, init_mode=1, init_rows=rows, max_iter=5)
    assert len(hist) == 2 and np.array_equal(hist[0], y) and np.array_equal(hist[1], y)
    hist0, _ = O.kshape(X, k, init_labels=y, max_iter=5)
    assert len(hist0) == 1


@pytest.mark.parametrize("init_mode", [0, 1])
def test_collapse_stop_matches_numpy_restatement(init_mode):
    """Two-class data fitted with k = 4: the clusters collapse to exactly
    k/2 = 2 non-empty ones.  With the GPU loop's collapse break the fit stops
    right after that refinement, keeping the labels of the assignment
    before it; without it the loop runs on (the CPU path)."""
    n, m, k = 60, 32, 4
    Xf, _ = T.generate(2, n, m, 2, 0.1)
    X = O.zscore(Xf.astype(np.float64))
    if init_mode == 0:
        init = T.init_labels(2, n, k)
        hist, _, col = O.kshape(X, k, init_labels=init, max_iter=10, seed=2, exact_steps=True,
                                stop_on_collapse=True, return_collapsed=True)
        ref, _ = np_kshape(X, k, init, 10, 2, stop_on_collapse=True)
    else:
        rows = T.init_rows(2, n, k)
        hist, _, col = O.kshape(X, k, init_mode=1, init_rows=rows, max_iter=10, seed=2,
                                exact_steps=True, stop_on_collapse=True, return_collapsed=True)
        ref, _ = np_kshape_tsl(X, k, rows, 10, 2, stop_on_collapse=True)
    assert col
    assert hist.shape == ref.shape and np.array_equal(hist, ref)
    assert len(set(hist[-1])) == k // 2
    full = O.kshape(X, k, init_mode=init_mode, max_iter=10, seed=2, exact_steps=True,
                    **({"init_labels": init} if init_mode == 0 else {"init_rows": rows}))[0]
    assert len(full) > len(hist)


def test_sum_pairwise_is_numpy_sum_bit_exact():
    rng = np.random.default_rng(17)
    for _ in range(400):
        m = int(rng.integers(1, 5000))
        a = rng.standard_normal(m) * 10.0 ** int(rng.integers(-3, 4))
        a -= a.mean()
        assert O.sum_pairwise(a) == a.sum()


def test_literal_skip_rule_fires_on_nonzero_centroids():
    """`cur_center.sum() != 0` (test_timesC.py:128) is also False for
    z-scored centroids whose float64 sum rounds to exactly 0.0: the
    reference then skips the alignment of that cluster.  Measured here on
    z-scored vectors; the engine and the default oracle skip only all-zero
    centroids."""
    rng = np.random.default_rng(1)
    hits = 0
    for _ in range(2000):
        c = np_zscore(rng.standard_normal(128) + np.sin(np.arange(128) / 7.0), 1)
        hits += O.sum_pairwise(c) == 0.0
    assert 0.02 < hits / 2000 < 0.3


def test_literal_skip_rule_on_an_exact_zero_sum_centroid():
    """A non-zero centroid whose NumPy sum is exactly 0.0: the literal rule
    leaves its members unaligned (like a zero centroid), the default rule
    aligns them."""
    n, m, k = 40, 24, 2
    Xf, _ = T.generate(8, n, m, k, 0.3)
    X = O.zscore(Xf.astype(np.float64))
    lab = (np.arange(n) % 2).astype(np.int32)
    t = np.arange(m)
    c = np.sin(2 * np.pi * t / m)
    c[m // 2:] = -c[:m // 2][::-1]  # exactly antisymmetric: pairwise sum == 0.0
    assert c.sum() == 0.0 and O.sum_pairwise(c) == 0.0 and np.any(c != 0)
    C = np.stack([c, c])
    lit, _, _ = O.refine(X, lab, C, seed=3, it=1, skip_rule=O.SKIP_NUMPY_SUM)
    unaligned, _, _ = O.refine(X, lab, np.zeros((k, m)), seed=3, it=1)
    aligned, _, _ = O.refine(X, lab, C, seed=3, it=1)
    assert np.array_equal(lit, unaligned)
    assert not np.allclose(aligned, unaligned)


def test_refine_with_given_shifts_equals_own_shifts():
    """orc_refine_shifts with every row's own alignment shift (against its
    cluster's old centroid) gives exactly orc_refine's result."""
    n, m, k = 50, 30, 3
    Xf, _ = T.generate(13, n, m, k, 0.3)
    X = O.zscore(Xf.astype(np.float64))
    lab = T.init_labels(13, n, k)
    C = O.zscore(X[T.init_rows(13, n, k)], 1)
    own = np.array([O.align_shift(X[i], C[lab[i]]) for i in range(n)], np.int32)
    a, _, _ = O.refine(X, lab, C, seed=1, it=1)
    b, _, _ = O.refine(X, lab, C, seed=1, it=1, shifts=own)
    assert np.array_equal(a, b)
    c, _, _ = O.refine(X, lab, C, seed=1, it=1, shifts=np.zeros(n, np.int32))
    assert not np.allclose(a, c)


def test_kshape_true_labels_are_a_fixed_point():
    """Noise-free, shift-free, well-separated classes: starting from the true
    labels the labels never change."""
    n, m, k = 30, 40, 3
    t = np.arange(m)
    protos = [np.sin(2 * np.pi * f * t / m) for f in (1, 3, 7)]
    y = np.repeat(np.arange(k), n // k).astype(np.int32)
    rng = np.random.default_rng(0)
    X = O.zscore(np.stack([rng.uniform(0.5, 2) * protos[c] + rng.normal() for c in y]))
    hist, _ = O.kshape(X, k, init_labels=y, max_iter=5)
    assert len(hist) == 1 and np.array_equal(hist[0], y)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("convention", [O.CONV_CPU, O.CONV_KSHAPE])
def test_kshape_converges_to_ground_truth(seed, convention):
    """Low-noise generator data (sigma=0.05): both alignment conventions
    recover the generating classes."""
    n, m, k = 90, 64, 3
    X, y = T.generate(seed, n, m, k, 0.05)
    Z = O.zscore(X.astype(np.float64))
    hist, _ = O.kshape(Z, k, init_labels=T.init_labels(seed, n, k), max_iter=30,
                       convention=convention)
    lab = hist[-1]
    # a perfect clustering up to label permutation
    for c in range(k):
        assert len(set(lab[y == c])) == 1
    assert len(set(lab)) == k
