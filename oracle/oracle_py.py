"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check one small run, and by bench.py's cpu_baseline leg.  The
product library never uses it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CONV_CPU, CONV_KSHAPE = 0, 1
SKIP_ZERO, SKIP_NUMPY_SUM = 0, 1


class ShapeInfo(ctypes.Structure):
    _fields_ = [
        ("lam", ctypes.c_double),
        ("steps", ctypes.c_int),
        ("fallback", ctypes.c_int),
        ("flipped", ctypes.c_int),
        ("empty", ctypes.c_int),
        ("sumdiff", ctypes.c_double),
    ]


_lib = None
_P = ctypes.c_void_p


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    I, L, D, U = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_uint64
    sig = {
        "orc_fft_size": (I, [I]),
        "orc_fft": (None, [_P, _P, I, I]),
        "orc_zscore": (None, [_P, L, I, I]),
        "orc_roll_zeropad": (None, [_P, I, I, _P]),
        "orc_ncc": (None, [_P, _P, I, _P]),
        "orc_align_shift": (I, [_P, _P, I, I]),
        "orc_assign": (None, [_P, L, I, _P, I, _P, _P, _P, _P, I]),
        "orc_power_v0": (None, [U, I, I, I, _P]),
        "orc_extract_shape": (I, [_P, L, I, _P, I, _P, ctypes.POINTER(ShapeInfo), I]),
        "orc_refine": (None, [_P, L, I, I, _P, _P, I, U, I, I, _P, _P, I, I, I]),
        "orc_refine_shifts": (None, [_P, L, I, I, _P, _P, I, U, I, I, _P, _P, I, I, I, _P]),
        "orc_kshape": (I, [_P, L, I, I, I, _P, _P, I, I, U, I, _P, _P, I, I, I, I, _P]),
        "orc_sum_pairwise": (D, [_P, L]),
        "orc_top_eig": (I, [_P, I, ctypes.POINTER(D), _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def fft_size(m):
    return load().orc_fft_size(m)


def fft(x, inverse=False):
    x = np.asarray(x, np.complex128)
    re, im = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    load().orc_fft(_p(re), _p(im), x.size, 1 if inverse else -1)
    return re + 1j * im


def zscore(x, ddof=0):
    x = _f64(x).copy()
    x2 = x.reshape(-1, x.shape[-1])
    load().orc_zscore(_p(x2), x2.shape[0], x2.shape[1], ddof)
    return x2.reshape(x.shape)


def roll_zeropad(a, shift):
    a = _f64(a)
    out = np.empty_like(a)
    load().orc_roll_zeropad(_p(a), a.size, shift, _p(out))
    return out


def ncc(x, y):
    x, y = _f64(x), _f64(y)
    out = np.empty(2 * x.size - 1)
    load().orc_ncc(_p(x), _p(y), x.size, _p(out))
    return out


def align_shift(member, centroid, convention=CONV_CPU):
    member, centroid = _f64(member), _f64(centroid)
    return load().orc_align_shift(_p(member), _p(centroid), member.size, convention)


def assign(X, C, dmat=False, nthreads=0):
    X, C = _f64(X), _f64(C)
    n, m = X.shape
    k = C.shape[0]
    lab = np.empty(n, np.int32)
    dist = np.empty(n)
    sh = np.empty(n, np.int32)
    dm = np.empty((n, k)) if dmat else None
    load().orc_assign(_p(X), n, m, _p(C), k, _p(lab), _p(dist), _p(sh), _p(dm), nthreads)
    return (lab, dist, sh, dm) if dmat else (lab, dist, sh)


def power_v0(seed, it, cluster, m):
    out = np.empty(m)
    load().orc_power_v0(seed, it, cluster, m, _p(out))
    return out


def extract_shape(A, v0, exact_steps=False, cal_select=0):
    A = _f64(A).reshape(-1, np.asarray(v0).size)
    v0 = _f64(v0)
    m = v0.size
    out = np.empty(m)
    info = ShapeInfo()
    rc = load().orc_extract_shape(_p(A), A.shape[0], m, _p(v0), int(exact_steps), _p(out),
                                  ctypes.byref(info), cal_select)
    if rc:
        raise RuntimeError(f"orc_extract_shape failed ({rc})")
    return out, {name: getattr(info, name) for name, _ in info._fields_}


def sum_pairwise(a):
    a = _f64(a)
    return load().orc_sum_pairwise(_p(a), a.size)


def refine(X, labels, C, convention=CONV_CPU, seed=0, it=0, exact_steps=False, nthreads=0,
           skip_rule=SKIP_ZERO, cal_select=0, shifts=None):
    """One refinement; shifts (n, optional): the alignment shift of every row
    instead of the one the oracle computes."""
    X = _f64(X)
    n, m = X.shape
    C = _f64(C).copy()
    k = C.shape[0]
    lab = np.ascontiguousarray(labels, np.int32)
    counts = np.empty(k, np.int32)
    infos = (ShapeInfo * k)()
    sh = None if shifts is None else np.ascontiguousarray(shifts, np.int32)
    load().orc_refine_shifts(_p(X), n, m, k, _p(lab), _p(C), convention, seed, it,
                             int(exact_steps), _p(counts), ctypes.cast(infos, ctypes.c_void_p),
                             nthreads, skip_rule, cal_select, _p(sh))
    return C, counts, [{name: getattr(i, name) for name, _ in i._fields_} for i in infos]


def kshape(X, k, init_mode=0, init_labels=None, init_rows=None, max_iter=10,
           convention=CONV_CPU, seed=0, exact_steps=False, nthreads=0, skip_rule=SKIP_ZERO,
           cal_select=0, stop_on_collapse=False, return_collapsed=False):
    X = _f64(X)
    n, m = X.shape
    hist = np.zeros((max_iter, n), np.int32)
    cent = np.zeros((k, m))
    il = None if init_labels is None else np.ascontiguousarray(init_labels, np.int32)
    ir = None if init_rows is None else np.ascontiguousarray(init_rows, np.int32)
    col = ctypes.c_int()
    it = load().orc_kshape(_p(X), n, m, k, init_mode, _p(il), _p(ir), max_iter, convention,
                           seed, int(exact_steps), _p(hist), _p(cent), nthreads, skip_rule,
                           cal_select, int(stop_on_collapse), ctypes.byref(col))
    if return_collapsed:
        return hist[:it], cent, bool(col.value)
    return hist[:it], cent


def top_eig(M):
    M = _f64(M)
    m = M.shape[0]
    lam = ctypes.c_double()
    vec = np.empty(m)
    rc = load().orc_top_eig(_p(M), m, ctypes.byref(lam), _p(vec))
    if rc:
        raise RuntimeError("orc_top_eig failed")
    return lam.value, vec
