"""The benchmark configurations end to end on the GPU, through the C ABI and
the drop-in CLI, against the oracle's committed golden fixtures
(tests/golden/*_oracle.npz, written by tests/golden/make_fixtures.py)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import timesc as T
from conftest import ROOT
from parity import CENT_RTOL, centroid_errors

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_fixture(name):
    f = np.load(os.path.join(GOLDEN, f"{name}_oracle.npz"), allow_pickle=False)
    return {k: f[k] for k in f.files}


def run_fit(ctx, fx):
    seed, n, m, k = int(fx["seed"]), int(fx["n"]), int(fx["m"]), int(fx["k"])
    frac = fx["frac"] if fx["frac"].size else None
    X, _ = T.generate(seed, n, m, k, float(fx["sigma"]), frac)
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        Z = d.numpy()
        kw = {}
        if int(fx["init_mode"]) == 0:
            kw["init_labels"] = T.init_labels(seed, n, k)
        else:
            kw["init_rows"] = T.init_rows(seed, n, k)
        res = ctx.fit(d, k, max_iter=int(fx["iters"]), init_mode=int(fx["init_mode"]), seed=seed,
                      labels_per_iter=True, **kw)
    finally:
        d.free()
    return X, Z, res


@pytest.mark.parametrize("name", ["cfg0", "cfg1"])
def test_config_matches_golden_fixture(ctx, name):
    fx = load_fixture(name)
    X, Z, res = run_fit(ctx, fx)
    got = res["labels_per_iter"]
    assert got.shape == fx["labels_per_iter"].shape, (got.shape, fx["labels_per_iter"].shape)
    assert np.array_equal(got, fx["labels_per_iter"])
    errs = centroid_errors(res["centroids"], fx["centroids"])
    assert max(errs) <= CENT_RTOL, errs


def test_cfg1_matches_live_oracle(ctx):
    """The GPU-normalized input through the live oracle (FP32 input rounding
    of the z-norm included) gives the same trajectory as the fixture."""
    fx = load_fixture("cfg1")
    X, Z, res = run_fit(ctx, fx)
    hist, cent = O.kshape(Z.astype(np.float64), int(fx["k"]), init_mode=1,
                          init_rows=T.init_rows(int(fx["seed"]), int(fx["n"]), int(fx["k"])),
                          max_iter=int(fx["iters"]), seed=int(fx["seed"]))
    assert np.array_equal(res["labels_per_iter"], hist)
    assert max(centroid_errors(res["centroids"], cent)) <= CENT_RTOL


@pytest.mark.parametrize("init_chose", [0, 1])
def test_cli_writes_the_fit_labels(ctx, tmp_path, init_chose):
    """`test -d 0 <file> k 0 <init_chose>` on a labeled CSV: one-based labels
    "l,l,...," equal to the fit through the C ABI (k from the labels, 10
    iterations, test_kshape.cu:21-30)."""
    X, y = T.generate(3, 300, 96, 3, 0.3)
    csv = str(tmp_path / "data.csv")
    T.write_labeled_csv(csv, X, y)
    out = str(tmp_path / "result.txt")
    dump = str(tmp_path / "iters.txt")
    r = subprocess.run([T.CLI_PATH, "-d", "0", csv, "99", "0", str(init_chose), "--out=" + out,
                        "--dump-iters=" + dump], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    text = open(out).read()
    assert text.endswith(",")
    cli = np.array([int(v) for v in text[:-1].split(",")])
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=0)
        res = ctx.fit(d, 3, max_iter=10, init_mode=init_chose, seed=0, labels_per_iter=True,
                      stop_on_collapse=True)
    finally:
        d.free()
    assert np.array_equal(cli, res["labels"] + 1)
    rows = [list(map(int, line.split(","))) for line in open(dump).read().split()]
    assert np.array_equal(np.array(rows), res["labels_per_iter"])


def test_cli_binary_input_and_cpu_label_format(ctx, tmp_path):
    """The same data as a binary matrix gives the same labels as the CSV;
    --cpu-format writes them zero-based, one per line."""
    X, y = T.generate(5, 240, 64, 3, 0.3)
    csv, binf = str(tmp_path / "d.csv"), str(tmp_path / "d.tcxb")
    T.write_labeled_csv(csv, X, y)
    T.write_binary(binf, X, y)
    outs = []
    for path, extra in ((csv, []), (binf, []), (binf, ["--cpu-format"])):
        out = str(tmp_path / f"r{len(outs)}.txt")
        r = subprocess.run([T.CLI_PATH, "-d", "0", path, "3", "0", "0", "--out=" + out] + extra,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(open(out).read())
    assert outs[0] == outs[1]
    one_based = np.array([int(v) for v in outs[0][:-1].split(",")])
    zero_based = np.array([int(v) for v in outs[2].split()])
    assert outs[2].endswith("\n") and np.array_equal(zero_based + 1, one_based)
