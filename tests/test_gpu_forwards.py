"""Runs the reference-name forwards of include/timesc_ref.hpp (z_norm_gpu_x,
NCC_3D, sbd3D, extract_shape, k_shape, k_shape_tsl, readFile,
read_unlabeled_File, writeFile) from a C++ caller (tests/ref_forwards.cpp)
and compares what such a caller gets with the C ABI on the same inputs.

The two alignment signs a reference caller meets:
  - sbd3D<float> keeps the reference GPU path's sign, NCC_3D(center, member)
    then zeroShift by argmax + 1 - m (SBD.cuh:61, 78): the classic k-Shape
    alignment, tcx_sbd_align(TCX_CONV_KSHAPE);
  - extract_shape<float>, k_shape<float> and k_shape_tsl<float> align with
    the CPU path's sign (test_timesC.py:129-132), the parity target."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import timesc as T

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def forwards(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("forwards")
    n, m, k = 120, 40, 3
    X, y = T.generate(12, n, m, k, 0.3)
    csv, ucsv = str(tmp / "l.csv"), str(tmp / "u.csv")
    T.write_labeled_csv(csv, X, y)
    with open(ucsv, "w") as f:  # unlabeled rows end with a comma (readfile.cuh:72-116)
        for row in X[:7]:
            f.write(",".join(f"{v:.9g}" for v in row) + ",\n")
    exe = str(tmp / "ref_forwards")
    libdir = os.path.dirname(T.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "ref_forwards.cpp"), "-o", exe,
                           "-L" + libdir, "-ltimesc_hip", "-Wl,-rpath," + libdir])
    out = tmp / "out"
    out.mkdir()
    print(f"\nrunning {exe}", flush=True)
    r = subprocess.run([exe, csv, ucsv, str(out), str(k)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ref_forwards done" in r.stdout

    def load(name, dtype):
        return np.fromfile(str(out / f"{name}.bin"), dtype=dtype)

    return dict(X=X, y=y, n=n, m=m, k=k, load=load, out=out, csv=csv)


def test_readers_and_writer(forwards):
    f = forwards
    X, y, n, m = f["X"], f["y"], f["n"], f["m"]
    vals = f["load"]("read_values", np.float32).reshape(n, m)
    assert np.array_equal(vals, T.read_labeled_csv(f["csv"])[0])
    assert np.array_equal(vals, X)  # %.9g round-trips FP32
    assert np.array_equal(f["load"]("read_labels", np.int32), y)
    assert np.array_equal(f["load"]("read_unlabeled", np.float32).reshape(7, m), X[:7])
    lab = f["load"]("kshape_tsl_labels", np.int32)
    assert open(str(f["out"] / "labels.txt")).read() == "".join(f"{v}," for v in lab)


def test_znorm_ncc_and_alignment(ctx, forwards):
    f = forwards
    X, n, m = f["X"], f["n"], f["m"]
    z = f["load"]("znorm", np.float32).reshape(n, m)
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=1)
        zc = d.numpy()
        assert np.array_equal(z, zc)
        assert np.abs(z - O.zscore(X.astype(np.float64), ddof=1)).max() <= 2e-6
        dx, dy = ctx.to_device(zc[:5]), ctx.to_device(zc[5:8])
        try:
            ncc = ctx.ncc(dx, dy)
        finally:
            dx.free()
            dy.free()
        assert np.array_equal(f["load"]("ncc", np.float32).reshape(5, 3, 2 * m - 1), ncc)
        dmem, dcen = ctx.to_device(zc[8:16]), ctx.to_device(zc[0])
        try:
            ctx.sbd_align(dmem, dcen, convention=T.CONV_KSHAPE)
            al = dmem.numpy()
        finally:
            dmem.free()
            dcen.free()
    finally:
        d.free()
    got = f["load"]("sbd3d", np.float32).reshape(8, m)
    assert np.array_equal(got, al)
    # the k-Shape sign: NCC(center, member), shift argmax + 1 - m
    for i in range(8):
        s = O.align_shift(zc[8 + i].astype(np.float64), zc[0].astype(np.float64), O.CONV_KSHAPE)
        assert np.array_equal(got[i], O.roll_zeropad(zc[8 + i], s).astype(np.float32))


def test_extract_shape_forward(ctx, forwards):
    f = forwards
    X, y, n, m, k = f["X"], f["y"], f["n"], f["m"], f["k"]
    d = ctx.to_device(X)
    lab0 = (y % k).astype(np.int32)
    try:
        ctx.znorm(d, ddof=1)
        zc = d.numpy()
        dl, dc = ctx.to_device(lab0), ctx.to_device(zc[:k].copy())
        try:
            sc = ctx.extract_shape(dl, d, None, dc, it=0)
            cent = dc.numpy()
        finally:
            dl.free()
            dc.free()
    finally:
        d.free()
    assert np.array_equal(f["load"]("extract_centroids", np.float32).reshape(k, m), cent)
    assert np.array_equal(f["load"]("extract_counts", np.int32), np.bincount(lab0, minlength=k))
    assert f["load"]("extract_sum_class", np.int32)[0] == sc == k
    # the CPU path's alignment sign, against the FP64 oracle
    ref, _, _ = O.refine(zc.astype(np.float64), lab0, zc[:k].astype(np.float64), seed=0, it=0)
    assert max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(cent, ref)) <= 1e-4


@pytest.mark.parametrize("name,init_mode", [("kshape", 0), ("kshape_tsl", 1)])
def test_kshape_forwards(ctx, forwards, name, init_mode):
    """k_shape / k_shape_tsl: one-based labels of the fit with seed 0, a cap
    of 10 iterations and the collapse stop; iter_num = iterations run when
    the labels stop changing."""
    f = forwards
    X, n, k = f["X"], f["n"], f["k"]
    d = ctx.to_device(X)
    try:
        ctx.znorm(d, ddof=1)
        res = ctx.fit(d, k, max_iter=10, init_mode=init_mode, seed=0, stop_on_collapse=True)
    finally:
        d.free()
    labels = f["load"](f"{name}_labels", np.int32)
    it, sc = f["load"](f"{name}_info", np.int32)
    assert np.array_equal(labels, res["labels"] + 1)
    assert sc == res["stats"]["sum_class"]
    assert it == (10 if res["stats"]["collapsed"] else res["stats"]["iterations"])
