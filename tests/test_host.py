"""Host-side checks of the engine library without a GPU: the C ABI exports
every symbol include/timesc.h declares, the synthetic generator, the CSV
readers/writer (format of Code/include/readfile.cuh), the drop-in CLI's
argument handling and the reference-name C++ shim."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import timesc as T
from conftest import ROOT, gpu_present

HEADER = os.path.join(ROOT, "include", "timesc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tcx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(T.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(T.EXPORTS) == names


def test_reference_shim_header_compiles(tmp_path):
    src = tmp_path / "shim.cpp"
    src.write_text(
        '#include "timesc_ref.hpp"\n'
        "int main(int argc, char**) {\n"
        "  if (argc > 5) {  // link every forward, run none\n"
        "    std::vector<int> lab; int r = 0, c = 0, it = 1, sc = 0;\n"
        "    float* h = readFile<float>(\"x\", lab, r, c);\n"
        "    z_norm_gpu_x<float>(h, r, c, 1);\n"
        "    k_shape<float>(h, lab.data(), r, c, 2, it, sc, 0);\n"
        "    k_shape_tsl<float>(h, h, lab.data(), r, c, 2, it, sc, 0);\n"
        "    NCC_3D<float>(h, h, h, 1, 1, c);\n"
        "    sbd3D<float>(h, h, h, 1, c);\n"
        "    extract_shape<float>(lab.data(), h, lab.data(), h, 2, r, r, c, sc, 0);\n"
        "    writeFile<int>(\"y\", lab.data(), r);\n"
        "    delete[] read_unlabeled_File<float>(\"x\", r, c);\n"
        "  }\n"
        "  return 0;\n"
        "}\n")
    exe = tmp_path / "shim"
    libdir = os.path.dirname(T.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-I" + os.path.join(ROOT, "include"), str(src),
                           "-o", str(exe), "-L" + libdir, "-ltimesc_hip",
                           "-Wl,-rpath," + libdir])
    assert subprocess.call([str(exe)]) == 0


# --------------------------------------------------------------- generator
def test_generator_is_deterministic_and_balanced():
    X1, y1 = T.generate(5, 400, 64, 4, 0.3)
    X2, y2 = T.generate(5, 400, 64, 4, 0.3)
    X3, _ = T.generate(6, 400, 64, 4, 0.3)
    assert X1.dtype == np.float32 and X1.shape == (400, 64)
    assert np.array_equal(X1, X2) and np.array_equal(y1, y2)
    assert not np.array_equal(X1, X3)
    assert np.array_equal(np.bincount(y1), [100] * 4)
    assert np.all(np.isfinite(X1))


def test_generator_class_fractions():
    _, y = T.generate(1, 1370, 32, 2, 0.3, class_frac=[0.6, 0.4])
    assert np.array_equal(np.bincount(y), [822, 548])


def test_generator_class_signal():
    """Series of one class correlate better with each other than across
    classes (after undoing the offset and scale)."""
    X, y = T.generate(2, 200, 128, 2, 0.0)
    Z = (X - X.mean(1, keepdims=True)) / X.std(1, keepdims=True)
    within = np.mean([Z[i] @ Z[j] for i in range(20) for j in range(20) if y[i] == y[j] and i != j])
    across = np.mean([Z[i] @ Z[j] for i in range(20) for j in range(20) if y[i] != y[j]])
    assert within > across


def test_initial_labels_and_rows():
    lab = T.init_labels(3, 1000, 7)
    assert lab.min() >= 0 and lab.max() < 7 and len(set(lab)) == 7
    assert np.array_equal(lab, T.init_labels(3, 1000, 7))
    rows = T.init_rows(3, 50, 10)
    assert len(set(rows)) == 10 and rows.min() >= 0 and rows.max() < 50


# --------------------------------------------------------------------- CSV
def test_labeled_csv_roundtrip(tmp_path):
    X, y = T.generate(4, 30, 17, 3, 0.3)
    p = str(tmp_path / "d.csv")
    T.write_labeled_csv(p, X, y)
    X2, y2 = T.read_labeled_csv(p)
    assert np.array_equal(X, X2) and np.array_equal(y, y2)


def test_unlabeled_csv_trailing_comma_and_crlf(tmp_path):
    p = tmp_path / "u.csv"
    p.write_text("1.5,2,3,\n-4,5e-1,6,\n")  # trailing comma ignored
    X = T.read_unlabeled_csv(str(p))
    assert np.array_equal(X, np.array([[1.5, 2, 3], [-4, 0.5, 6]], np.float32))
    p.write_text("1.5,2,3\r\n-4,5e-1,6\r\n")  # "3\r" parses as 3 (atof)
    assert np.array_equal(T.read_unlabeled_csv(str(p)), X)
    # CRLF after a trailing comma: getline(ss, s, ',') yields a last field
    # "\r" that atof reads as 0, as readfile.cuh:87-93 does
    p.write_text("1.5,2,3,\r\n-4,5e-1,6,\r\n")
    X4 = T.read_unlabeled_csv(str(p))
    assert X4.shape == (2, 4) and np.all(X4[:, 3] == 0) and np.array_equal(X4[:, :3], X)


def test_label_parsing_like_stoi(tmp_path):
    p = tmp_path / "l.csv"
    p.write_text("1.0,0.5,0.25\n-1,1,2\n")  # stoi("1.0") == 1
    X, y = T.read_labeled_csv(str(p))
    assert list(y) == [1, -1] and X.shape == (2, 2)


@pytest.mark.parametrize("content", ["1,2,3\n4,5\n", "a,1,2\n", ""])
def test_csv_errors(tmp_path, content):
    p = tmp_path / "bad.csv"
    p.write_text(content)
    with pytest.raises(T.TimescError) as e:
        T.read_labeled_csv(str(p))
    assert e.value.code == T.ERR_IO


def test_missing_file_is_an_error():
    with pytest.raises(T.TimescError) as e:
        T.read_unlabeled_csv("/nonexistent/file.csv")
    assert e.value.code == T.ERR_IO


def test_value_parsing_like_atof(tmp_path):
    """Each value as atof then float: leading blanks, '+', exponents,
    inf/nan, hex floats, a valid prefix before junk, no digits -> 0,
    out-of-range magnitudes."""
    fields = ["1.5", "  -2.25", "+3", "1e-3", "6.02E23", "inf", "-Infinity", "0x1p3", "-0x1.8p1",
              "3.5abc", "abc", "", "-", "+-1", "1e400", "-1e400", "1e-400", "0.1", "-0",
              "123456789.123456789"]
    expect = [1.5, -2.25, 3.0, 1e-3, 6.02e23, np.inf, -np.inf, 8.0, -3.0, 3.5, 0.0, 0.0, 0.0, 0.0,
              np.inf, -np.inf, 0.0, 0.1, -0.0, 123456789.123456789]
    p = tmp_path / "v.csv"
    p.write_text("7," + ",".join(fields) + "\n" + "8," + ",".join(["nan"] * len(fields)) + "\n")
    X, y = T.read_labeled_csv(str(p))
    assert list(y) == [7, 8]
    assert np.array_equal(X[0], np.array(expect, np.float32))
    assert np.signbit(X[0][18]) and np.all(np.isnan(X[1]))


def test_large_csv_parsed_in_parallel_chunks(tmp_path):
    """A file of several 4 MB chunks: values and labels equal a plain
    sequential parse (Python float -> float32), and a ragged line in a later
    chunk is reported with its file line number."""
    rng = np.random.default_rng(1)
    n, m = 9000, 120
    X = rng.standard_normal((n, m)) * 10.0 ** rng.integers(-3, 4, (n, 1))
    y = rng.integers(-5, 100, n)
    lines = [f"{y[i]}," + ",".join(repr(float(v)) for v in X[i]) for i in range(n)]
    p = tmp_path / "big.csv"
    p.write_text("\n".join(lines) + "\n")
    assert p.stat().st_size > 3 * (4 << 20)
    got, lab = T.read_labeled_csv(str(p))
    assert np.array_equal(lab, y.astype(np.int32))
    assert np.array_equal(got, X.astype(np.float32))
    lines[7000] = lines[7000] + ",1.0"
    p.write_text("\n".join(lines) + "\n")
    with pytest.raises(T.TimescError) as e:
        T.read_labeled_csv(str(p))
    assert e.value.code == T.ERR_IO and "line 7001" in str(e.value)


def test_binary_matrix_roundtrip_and_errors(tmp_path):
    X, y = T.generate(2, 50, 33, 3, 0.3)
    p = str(tmp_path / "d.tcxb")
    T.write_binary(p, X, y)
    X2, y2 = T.read_binary(p)
    assert np.array_equal(X, X2) and np.array_equal(y, y2)
    T.write_binary(p, X)
    X3, y3 = T.read_binary(p)
    assert np.array_equal(X, X3) and y3 is None
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-4])
    with pytest.raises(T.TimescError) as e:
        T.read_binary(p)
    assert e.value.code == T.ERR_IO and "truncated" in str(e.value)
    open(p, "wb").write(b"NOTTCXB!" + raw[8:])
    with pytest.raises(T.TimescError):
        T.read_binary(p)


def test_write_labels_lines_format(tmp_path):
    """The CPU path's output: zero-based, one per line (np.savetxt fmt %.0f)."""
    p = str(tmp_path / "t.txt")
    T.write_labels_lines(p, [0, 3, 11])
    assert open(p).read() == "0\n3\n11\n"
    assert np.array_equal(np.loadtxt(p), [0, 3, 11])


def test_write_labels_format(tmp_path):
    p = str(tmp_path / "r.txt")
    T.write_labels(p, [1, 2, 10])
    assert open(p).read() == "1,2,10,"


# --------------------------------------------------------------------- CLI
def run_cli(*args):
    return subprocess.run([T.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


def test_cli_without_dash_d_exits_silently():
    r = run_cli("x")
    assert r.returncode == 0 and r.stdout == ""


def test_cli_usage_on_missing_arguments():
    r = run_cli("-d", "0", "file")
    assert r.returncode == 2 and "usage" in r.stderr


def test_cli_reports_a_missing_file():
    r = run_cli("-d", "0", "/nonexistent.csv", "3", "0", "0")
    assert r.returncode == 1 and "cannot read file" in r.stderr


@pytest.mark.skipif(gpu_present(), reason="checks the error path of a host without a GPU")
def test_context_creation_fails_cleanly_without_gpu():
    with pytest.raises(T.TimescError) as e:
        T.Context(0)
    assert e.value.code in (T.ERR_ARG, T.ERR_HIP)


def test_tune_table_roundtrip_and_validation(tmp_path):
    """The autotuner's table (host only): entries set, saved, reset, loaded
    back; malformed files and invalid entries are refused and leave the
    table unchanged."""
    T.tune_reset()
    try:
        T.tune_set(T.TUNE_PRUNE_CAP, 1024, 1, 16, 4096)
        T.tune_set(T.TUNE_PACK_ROWS, 512, 0, 0, 16)
        assert T.tune_get(T.TUNE_PRUNE_CAP, 1024, 1, 16) == 4096
        assert T.tune_get(T.TUNE_PACK_ROWS, 500) == 16  # m rounded up to 32
        assert T.tune_get(T.TUNE_PRUNE_CAP, 2048, 1, 16) == 0  # no entry: built-in
        assert T.tune_get(T.TUNE_PRUNE_CAP, 1024, 1, 11) == 0  # another size bucket
        path = tmp_path / "tune.txt"
        T.tune_save(path)
        text = path.read_text()
        assert "prune 1024 1 16 4096" in text and "pack 512 16" in text
        T.tune_reset()
        assert T.tune_get(T.TUNE_PRUNE_CAP, 1024, 1, 16) == 0
        T.tune_load(path)
        assert T.tune_get(T.TUNE_PRUNE_CAP, 1024, 1, 16) == 4096
        assert T.tune_get(T.TUNE_PACK_ROWS, 512) == 16
        for bad in ["prune 1000 1 16 4096\n",     # F not a power of two
                    "pack 512 0 24\n",             # rows not 16/32/64
                    "pack 1024 64\n",              # LDS of 64 rows at m = 1024 too large
                    "pack 2048 16\n",              # m beyond the LDS form
                    "prune 1024 1 16\n", "tile 3 4 5\n"]:
            p = tmp_path / "bad.txt"
            p.write_text("prune 4096 2 12 2048\n" + bad)
            with pytest.raises(T.TimescError):
                T.tune_load(p)
            assert T.tune_get(T.TUNE_PRUNE_CAP, 4096, 2, 12) == 0  # nothing applied
        with pytest.raises(T.TimescError):
            T.tune_set(T.TUNE_PACK_ROWS, 100, 0, 0, 16)  # key not a multiple of 32
        with pytest.raises(T.TimescError):
            T.tune_load(tmp_path / "missing.txt")
    finally:
        T.tune_reset()
