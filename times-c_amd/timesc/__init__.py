"""ctypes binding of libtimesc_hip.so (the engine's C ABI, include/timesc.h).

Used by the tests, bench.py and __graft_entry__.smoke().  Device memory is
managed through the library's own helpers (tcx_malloc / tcx_memcpy_*), so no
other GPU runtime is loaded into the process.  Loading fails loudly if the
library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes
import itertools
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libtimesc_hip.so")
CLI_PATH = os.path.join(PKG_ROOT, "bin", "test")

OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_BUDGET, ERR_IO, ERR_UNSUPPORTED, ERR_NUMERIC = range(8)
CONV_CPU, CONV_KSHAPE = 0, 1
S_SPLIT_BF16, S_BF16, S_AUTO = 0, 1, 2

# every symbol include/timesc.h declares
EXPORTS = [
    "tcx_last_error", "tcx_create", "tcx_destroy", "tcx_workspace_bytes",
    "tcx_fit_workspace_bytes", "tcx_malloc",
    "tcx_free", "tcx_memcpy_h2d", "tcx_memcpy_d2h", "tcx_memset", "tcx_synchronize",
    "tcx_znorm", "tcx_ncc", "tcx_sbd_align", "tcx_assign", "tcx_assign_noalloc",
    "tcx_extract_shape",
    "tcx_fit_opts_default", "tcx_kshape_fit", "tcx_fit_begin", "tcx_fit_step", "tcx_fit_end",
    "tcx_fit_query", "tcx_fit_snapshot",
    "tcx_kshape", "tcx_kshape_tsl", "tcx_batch_create", "tcx_batch_run",
    "tcx_batch_run_grouped", "tcx_batch_destroy",
    "tcx_server_create", "tcx_server_add_model", "tcx_server_submit", "tcx_server_wait",
    "tcx_server_destroy",
    "tcx_read_labeled_csv", "tcx_read_unlabeled_csv", "tcx_write_labels",
    "tcx_write_labels_lines", "tcx_write_binary", "tcx_read_binary", "tcx_free_host",
    "tcx_generate", "tcx_init_labels", "tcx_init_rows", "tcx_power_v0",
    "tcx_autotune", "tcx_tune_set", "tcx_tune_get", "tcx_tune_reset", "tcx_tune_save",
    "tcx_tune_load",
]

TUNE_PRUNE_CAP, TUNE_PACK_ROWS = 0, 1


class TuneReport(ctypes.Structure):
    """tcx_tune_report."""
    _fields_ = [
        ("fft", ctypes.c_int),
        ("prune_cap", ctypes.c_int),
        ("pack_rows", ctypes.c_int),
        ("ms_default", ctypes.c_double),
        ("ms_tuned", ctypes.c_double),
    ]


class TimescError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"timesc status {code}: {msg}")
        self.code = code


class FitOpts(ctypes.Structure):
    _fields_ = [
        ("max_iter", ctypes.c_int),
        ("init_mode", ctypes.c_int),
        ("seed", ctypes.c_uint64),
        ("init_labels", ctypes.POINTER(ctypes.c_int)),
        ("init_rows", ctypes.POINTER(ctypes.c_int)),
        ("convention", ctypes.c_int),
        ("s_precision", ctypes.c_int),
        ("cal_select", ctypes.c_int),
        ("power_tol", ctypes.c_double),
        ("labels_per_iter", ctypes.POINTER(ctypes.c_int)),
        ("stop_on_collapse", ctypes.c_int),
        ("s_budget", ctypes.c_double),
    ]


class FitStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("sum_class", ctypes.c_int),
        ("power_steps_max", ctypes.c_int),
        ("fallbacks", ctypes.c_int),
        ("ms_total", ctypes.c_double),
        ("ms_refine", ctypes.c_double),
        ("ms_assign", ctypes.c_double),
        ("sbd_evals", ctypes.c_longlong),
        ("sbd_pairs", ctypes.c_longlong),
        ("ms_assign_bounds", ctypes.c_double),
        ("collapsed", ctypes.c_int),
        ("s_bf16_from", ctypes.c_int),
        ("s_check_err", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


REQ_FIT, REQ_ASSIGN = 0, 1


class Request(ctypes.Structure):
    """tcx_request: one fit or assignment query of a server."""
    _fields_ = [
        ("kind", ctypes.c_int),
        ("h_x", ctypes.c_void_p),
        ("n", ctypes.c_int),
        ("m", ctypes.c_int),
        ("znorm_ddof", ctypes.c_int),
        ("k", ctypes.c_int),
        ("opts", FitOpts),
        ("publish_model", ctypes.c_int),
        ("model", ctypes.c_int),
        ("h_labels", ctypes.c_void_p),
        ("h_centroids", ctypes.c_void_p),
        ("h_dist", ctypes.c_void_p),
        ("h_shift", ctypes.c_void_p),
        ("model_out", ctypes.c_int),
        ("stats", FitStats),
        ("status", ctypes.c_int),
        ("done", ctypes.c_int),
        ("ms_queued", ctypes.c_double),
        ("ms_run", ctypes.c_double),
        ("ms_latency", ctypes.c_double),
        ("message", ctypes.c_char * 160),
    ]


_lib = None

# Every open Context, Batch and Server, closed at interpreter exit (servers
# and batches first) so that no device context outlives the runtime.
_live = weakref.WeakValueDictionary()
_live_seq = itertools.count()


def _register(obj, rank):
    _live[(rank, next(_live_seq))] = obj


@atexit.register
def _close_all():
    for key in sorted(list(_live.keys())):
        obj = _live.get(key)
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass


_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_long
_Z = ctypes.c_size_t

ORDER_GIVEN, ORDER_SHORTEST_FIRST, ORDER_LONGEST_FIRST = 0, 1, 2


class Job(ctypes.Structure):
    """tcx_job: one fit of a batch."""
    _fields_ = [
        ("d_x", ctypes.c_void_p),
        ("n", ctypes.c_int),
        ("m", ctypes.c_int),
        ("k", ctypes.c_int),
        ("znorm_ddof", ctypes.c_int),
        ("opts", FitOpts),
        ("h_labels", ctypes.c_void_p),
        ("h_centroids", ctypes.c_void_p),
        ("status", ctypes.c_int),
        ("worker", ctypes.c_int),
        ("stats", FitStats),
        ("ms_queued", ctypes.c_double),
        ("ms_run", ctypes.c_double),
        ("ms_latency", ctypes.c_double),
        ("message", ctypes.c_char * 160),
    ]


def load():
    """Load the engine library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "tcx_last_error": (ctypes.c_char_p, []),
        "tcx_create": (_I, [_I, _Z, ctypes.POINTER(_P)]),
        "tcx_destroy": (_I, [_P]),
        "tcx_workspace_bytes": (_Z, [_P]),
        "tcx_fit_workspace_bytes": (_Z, [_P, _I, _I, _I]),
        "tcx_malloc": (_I, [_P, ctypes.POINTER(_P), _Z]),
        "tcx_free": (_I, [_P, _P]),
        "tcx_memcpy_h2d": (_I, [_P, _P, _P, _Z]),
        "tcx_memcpy_d2h": (_I, [_P, _P, _P, _Z]),
        "tcx_memset": (_I, [_P, _P, _I, _Z]),
        "tcx_synchronize": (_I, [_P]),
        "tcx_znorm": (_I, [_P, _P, _I, _I, _I]),
        "tcx_ncc": (_I, [_P, _P, _P, _P, _I, _I, _I]),
        "tcx_sbd_align": (_I, [_P, _P, _P, _I, _I, _I]),
        "tcx_assign": (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
        "tcx_assign_noalloc": (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
        "tcx_extract_shape": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, ctypes.POINTER(_I)]),
        "tcx_fit_opts_default": (None, [ctypes.POINTER(FitOpts)]),
        "tcx_kshape_fit": (_I, [_P, _P, _I, _I, _I, ctypes.POINTER(FitOpts), _P, _P,
                                ctypes.POINTER(FitStats)]),
        "tcx_fit_begin": (_I, [_P, _P, _I, _I, _I, ctypes.POINTER(FitOpts)]),
        "tcx_fit_step": (_I, [_P, ctypes.POINTER(_I)]),
        "tcx_fit_end": (_I, [_P, _P, _P, ctypes.POINTER(FitStats)]),
        "tcx_fit_query": (_I, [_P, ctypes.POINTER(FitStats)]),
        "tcx_fit_snapshot": (_I, [_P, _P, _P, _P, _P]),
        "tcx_kshape": (_I, [_P, _P, _P, _I, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), _I]),
        "tcx_kshape_tsl": (_I, [_P, _P, _P, _P, _I, _I, _I, ctypes.POINTER(_I),
                                ctypes.POINTER(_I), _I]),
        "tcx_batch_create": (_I, [_I, _I, _Z, ctypes.POINTER(_P)]),
        "tcx_batch_run": (_I, [_P, ctypes.POINTER(Job), _I, _I]),
        "tcx_batch_run_grouped": (_I, [_P, ctypes.POINTER(Job), _I]),
        "tcx_batch_destroy": (_I, [_P]),
        "tcx_server_create": (_I, [_I, _I, _I, _Z, ctypes.POINTER(_P)]),
        "tcx_server_add_model": (_I, [_P, _P, _I, _I, ctypes.POINTER(_I)]),
        "tcx_server_submit": (_I, [_P, ctypes.POINTER(Request)]),
        "tcx_server_wait": (_I, [_P, ctypes.POINTER(Request)]),
        "tcx_server_destroy": (_I, [_P]),
        "tcx_read_labeled_csv": (_I, [ctypes.c_char_p, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                      ctypes.POINTER(_I), ctypes.POINTER(_I)]),
        "tcx_read_unlabeled_csv": (_I, [ctypes.c_char_p, ctypes.POINTER(_P),
                                        ctypes.POINTER(_I), ctypes.POINTER(_I)]),
        "tcx_write_labels": (_I, [ctypes.c_char_p, _P, _I]),
        "tcx_write_labels_lines": (_I, [ctypes.c_char_p, _P, _I]),
        "tcx_write_binary": (_I, [ctypes.c_char_p, _P, _P, _I, _I]),
        "tcx_read_binary": (_I, [ctypes.c_char_p, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                 ctypes.POINTER(_I), ctypes.POINTER(_I)]),
        "tcx_free_host": (None, [_P]),
        "tcx_generate": (_I, [ctypes.c_uint64, _L, _I, _I, ctypes.c_double, _P, _P, _P]),
        "tcx_init_labels": (_I, [ctypes.c_uint64, _L, _I, _P]),
        "tcx_init_rows": (_I, [ctypes.c_uint64, _L, _I, _P]),
        "tcx_power_v0": (None, [ctypes.c_uint64, _I, _I, _I, _P]),
        "tcx_autotune": (_I, [_P, _P, _I, _I, _I, _I, ctypes.POINTER(TuneReport)]),
        "tcx_tune_set": (_I, [_I, _I, _I, _I, _I]),
        "tcx_tune_get": (_I, [_I, _I, _I, _I, ctypes.POINTER(_I)]),
        "tcx_tune_reset": (None, []),
        "tcx_tune_save": (_I, [ctypes.c_char_p]),
        "tcx_tune_load": (_I, [ctypes.c_char_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def check(rc):
    if rc != OK:
        raise TimescError(rc, load().tcx_last_error().decode())


# ------------------------------------------------------------- host utilities
def generate(seed, n, m, K, sigma, class_frac=None):
    """Seeded synthetic dataset: (X float32 [n,m], y int32 [n])."""
    lib = load()
    X = np.empty((n, m), np.float32)
    y = np.empty(n, np.int32)
    cf = None if class_frac is None else np.ascontiguousarray(class_frac, np.float64)
    check(lib.tcx_generate(seed, n, m, K, sigma, _ptr(cf), _ptr(X), _ptr(y)))
    return X, y


def init_labels(seed, n, k):
    out = np.empty(n, np.int32)
    check(load().tcx_init_labels(seed, n, k, _ptr(out)))
    return out


def init_rows(seed, n, k):
    out = np.empty(k, np.int32)
    check(load().tcx_init_rows(seed, n, k, _ptr(out)))
    return out


def power_v0(seed, it, cluster, m):
    out = np.empty(m, np.float64)
    load().tcx_power_v0(seed, it, cluster, m, _ptr(out))
    return out


def read_labeled_csv(path):
    lib = load()
    d, l = ctypes.c_void_p(), ctypes.c_void_p()
    r, c = ctypes.c_int(), ctypes.c_int()
    check(lib.tcx_read_labeled_csv(path.encode(), ctypes.byref(d), ctypes.byref(l),
                                   ctypes.byref(r), ctypes.byref(c)))
    try:
        X = np.ctypeslib.as_array((ctypes.c_float * (r.value * c.value)).from_address(d.value))
        y = np.ctypeslib.as_array((ctypes.c_int * r.value).from_address(l.value))
        return X.reshape(r.value, c.value).copy(), y.copy()
    finally:
        lib.tcx_free_host(d)
        lib.tcx_free_host(l)


def write_labels_lines(path, labels):
    lab = np.ascontiguousarray(labels, np.int32)
    check(load().tcx_write_labels_lines(path.encode(), _ptr(lab), lab.size))


def write_binary(path, X, labels=None):
    X = np.ascontiguousarray(X, np.float32)
    lab = None if labels is None else np.ascontiguousarray(labels, np.int32)
    check(load().tcx_write_binary(path.encode(), _ptr(X), _ptr(lab), X.shape[0], X.shape[1]))


def read_binary(path):
    """(X float32 [rows, cols], labels int32 [rows] or None)."""
    lib = load()
    d, l = ctypes.c_void_p(), ctypes.c_void_p()
    r, c = ctypes.c_int(), ctypes.c_int()
    check(lib.tcx_read_binary(path.encode(), ctypes.byref(d), ctypes.byref(l), ctypes.byref(r),
                              ctypes.byref(c)))
    try:
        X = np.ctypeslib.as_array((ctypes.c_float * (r.value * c.value)).from_address(d.value))
        X = X.reshape(r.value, c.value).copy()
        y = None
        if l.value:
            y = np.ctypeslib.as_array((ctypes.c_int * r.value).from_address(l.value)).copy()
        return X, y
    finally:
        lib.tcx_free_host(d)
        if l.value:
            lib.tcx_free_host(l)


def read_unlabeled_csv(path):
    lib = load()
    d = ctypes.c_void_p()
    r, c = ctypes.c_int(), ctypes.c_int()
    check(lib.tcx_read_unlabeled_csv(path.encode(), ctypes.byref(d), ctypes.byref(r),
                                     ctypes.byref(c)))
    try:
        X = np.ctypeslib.as_array((ctypes.c_float * (r.value * c.value)).from_address(d.value))
        return X.reshape(r.value, c.value).copy()
    finally:
        lib.tcx_free_host(d)


def write_labels(path, labels):
    lab = np.ascontiguousarray(labels, np.int32)
    check(load().tcx_write_labels(path.encode(), _ptr(lab), lab.size))


def write_labeled_csv(path, X, y):
    """Labeled CSV in the reference's input format; values as %.9g of the
    float32 value so every reader sees the same numbers."""
    X = np.asarray(X, np.float32)
    with open(path, "w") as f:
        for i in range(X.shape[0]):
            f.write(str(int(y[i])) + "," + ",".join("%.9g" % v for v in X[i]) + "\n")


# --------------------------------------------------------------- device side
class DeviceArray:
    """A device buffer owned by a Context."""

    def __init__(self, ctx, ptr, shape, dtype):
        self.ctx, self.ptr, self.shape, self.dtype = ctx, ptr, tuple(shape), np.dtype(dtype)

    @property
    def nbytes(self):
        return int(np.prod(self.shape)) * self.dtype.itemsize

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx.d2h(out, self)
        return out

    def free(self):
        if self.ptr:
            self.ctx.free(self)


def tune_set(kind, key0, key1, key2, value):
    """One table entry: (TUNE_PRUNE_CAP, F, ceil(k/16), floor(log2 n), cap)
    or (TUNE_PACK_ROWS, m rounded up to 32, 0, 0, rows)."""
    check(load().tcx_tune_set(kind, key0, key1, key2, value))


def tune_get(kind, key0, key1=0, key2=0):
    v = ctypes.c_int()
    check(load().tcx_tune_get(kind, key0, key1, key2, ctypes.byref(v)))
    return v.value


def tune_reset():
    load().tcx_tune_reset()


def tune_save(path):
    check(load().tcx_tune_save(str(path).encode()))


def tune_load(path):
    check(load().tcx_tune_load(str(path).encode()))


class Context:
    def __init__(self, device=0, budget=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.tcx_create(device, budget, ctypes.byref(h)))
        self.h = h
        _register(self, 2)

    def close(self):
        if self.h:
            check(self.lib.tcx_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def workspace_bytes(self):
        return self.lib.tcx_workspace_bytes(self.h)

    def fit_workspace_bound(self, n, m, k):
        """Upper bound of the workspace bytes one fit of (n, m, k) grows."""
        return self.lib.tcx_fit_workspace_bytes(self.h, n, m, k)

    def empty(self, shape, dtype):
        shape = (shape,) if np.isscalar(shape) else tuple(shape)
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = ctypes.c_void_p()
        check(self.lib.tcx_malloc(self.h, ctypes.byref(p), nbytes))
        return DeviceArray(self, p.value, shape, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = self.empty(a.shape, a.dtype)
        check(self.lib.tcx_memcpy_h2d(self.h, d.ptr, _ptr(a), a.nbytes))
        return d

    def h2d(self, d, a):
        a = np.ascontiguousarray(a, d.dtype)
        assert a.nbytes == d.nbytes
        check(self.lib.tcx_memcpy_h2d(self.h, d.ptr, _ptr(a), a.nbytes))

    def d2h(self, out, d):
        assert out.flags.c_contiguous and out.nbytes == d.nbytes
        check(self.lib.tcx_memcpy_d2h(self.h, _ptr(out), d.ptr, d.nbytes))

    def free(self, d):
        check(self.lib.tcx_free(self.h, d.ptr))
        d.ptr = None

    def sync(self):
        check(self.lib.tcx_synchronize(self.h))

    # primitives --------------------------------------------------------
    def znorm(self, d, ddof=0):
        rows, cols = d.shape
        check(self.lib.tcx_znorm(self.h, d.ptr, rows, cols, ddof))

    def ncc(self, dx, dy):
        xr, m = dx.shape
        yr = dy.shape[0]
        out = self.empty((xr, yr, 2 * m - 1), np.float32)
        try:
            check(self.lib.tcx_ncc(self.h, dx.ptr, dy.ptr, out.ptr, xr, yr, m))
            return out.numpy()
        finally:
            out.free()

    def sbd_align(self, dmembers, dcenter, convention=CONV_CPU):
        rows, m = dmembers.shape
        check(self.lib.tcx_sbd_align(self.h, dmembers.ptr, dcenter.ptr, rows, m, convention))

    def assign(self, dx, dc, dmat=False, noalloc=False):
        """tcx_assign (noalloc: tcx_assign_noalloc) of the rows of dx against
        the centroids dc; returns labels, distances, lags (and the distance
        matrix with dmat)."""
        n, m = dx.shape
        k = dc.shape[0]
        lab = self.empty(n, np.int32)
        dist = self.empty(n, np.float32)
        sh = self.empty(n, np.int32)
        dm = self.empty((n, k), np.float32) if dmat else None
        fn = self.lib.tcx_assign_noalloc if noalloc else self.lib.tcx_assign
        try:
            check(fn(self.h, dx.ptr, dc.ptr, n, m, k, lab.ptr, dist.ptr, sh.ptr,
                     dm.ptr if dm else None))
            out = (lab.numpy(), dist.numpy(), sh.numpy())
            if dm:
                out = out + (dm.numpy(),)
            return out
        finally:
            for a in (lab, dist, sh, dm):
                if a is not None:
                    a.free()

    def extract_shape(self, dlabel, dx, dshift, dcent, it=0, cal_select=0):
        n, m = dx.shape
        k = dcent.shape[0]
        sc = ctypes.c_int()
        check(self.lib.tcx_extract_shape(self.h, dlabel.ptr, dx.ptr,
                                         dshift.ptr if dshift is not None else None,
                                         dcent.ptr, n, m, k, it, cal_select, ctypes.byref(sc)))
        return sc.value

    def _opts(self, n, max_iter=10, init_mode=0, seed=0, init_labels=None, init_rows=None,
              convention=CONV_CPU, s_precision=S_SPLIT_BF16, power_tol=1e-12, cal_select=0,
              stop_on_collapse=False, labels_per_iter=False, s_budget=0.0):
        """FitOpts and the host arrays it points to (kept alive by the caller)."""
        o = FitOpts()
        self.lib.tcx_fit_opts_default(ctypes.byref(o))
        o.max_iter, o.init_mode, o.seed = max_iter, init_mode, seed
        o.cal_select = cal_select
        o.convention, o.s_precision, o.power_tol = convention, s_precision, power_tol
        o.stop_on_collapse = int(stop_on_collapse)
        o.s_budget = s_budget
        keep = {}
        if init_labels is not None:
            keep["il"] = np.ascontiguousarray(init_labels, np.int32)
            o.init_labels = keep["il"].ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        if init_rows is not None:
            keep["ir"] = np.ascontiguousarray(init_rows, np.int32)
            o.init_rows = keep["ir"].ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        if labels_per_iter:
            keep["hist"] = np.zeros((max_iter, n), np.int32)
            o.labels_per_iter = keep["hist"].ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        return o, keep

    def fit_begin(self, dx, k, **kw):
        """Starts a stepwise fit (tcx_fit_begin); keyword options as for fit()."""
        n, m = dx.shape
        o, keep = self._opts(n, **kw)
        self._fit_keep = keep  # labels_per_iter is written by every step
        check(self.lib.tcx_fit_begin(self.h, dx.ptr, n, m, k, ctypes.byref(o)))

    def fit_step(self):
        ch = ctypes.c_int()
        check(self.lib.tcx_fit_step(self.h, ctypes.byref(ch)))
        return ch.value

    def fit_query(self):
        st = FitStats()
        check(self.lib.tcx_fit_query(self.h, ctypes.byref(st)))
        return st.as_dict()

    def fit_snapshot(self, n, k, m):
        """Labels, distances and lags of the fit's last assignment and its
        current centroids."""
        labels = np.empty(n, np.int32)
        dist = np.empty(n, np.float32)
        shift = np.empty(n, np.int32)
        cent = np.empty((k, m), np.float32)
        check(self.lib.tcx_fit_snapshot(self.h, _ptr(labels), _ptr(dist), _ptr(shift),
                                        _ptr(cent)))
        return labels, dist, shift, cent

    def fit_end(self, n, k, m):
        labels = np.empty(n, np.int32)
        cent = np.empty((k, m), np.float32)
        st = FitStats()
        check(self.lib.tcx_fit_end(self.h, _ptr(labels), _ptr(cent), ctypes.byref(st)))
        return {"labels": labels, "centroids": cent, "stats": st.as_dict()}

    def autotune(self, dx, k, reps=3):
        """tcx_autotune on the z-normalized series dx: the launch shapes that
        do not change results, timed on 3-iteration fits (best of reps)."""
        n, m = dx.shape
        r = TuneReport()
        check(self.lib.tcx_autotune(self.h, dx.ptr, n, m, k, reps, ctypes.byref(r)))
        return {name: getattr(r, name) for name, _ in r._fields_}

    def fit(self, dx, k, **kw):
        """Full fit (tcx_kshape_fit).  Options: max_iter, init_mode, seed,
        init_labels, init_rows, convention, s_precision, power_tol, cal_select,
        stop_on_collapse, labels_per_iter (also return the labels after every
        iteration)."""
        n, m = dx.shape
        o, keep = self._opts(n, **kw)
        labels = np.empty(n, np.int32)
        cent = np.empty((k, m), np.float32)
        st = FitStats()
        check(self.lib.tcx_kshape_fit(self.h, dx.ptr, n, m, k, ctypes.byref(o), _ptr(labels),
                                      _ptr(cent), ctypes.byref(st)))
        res = {"labels": labels, "centroids": cent, "stats": st.as_dict()}
        if "hist" in keep:
            res["labels_per_iter"] = keep["hist"][: st.iterations]
        return res


class Batch:
    """A pool of worker contexts running independent fits concurrently."""

    def __init__(self, device=0, workers=4, budget=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.tcx_batch_create(device, workers, budget, ctypes.byref(h)))
        self.h = h
        _register(self, 1)

    def close(self):
        if self.h:
            check(self.lib.tcx_batch_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, specs, order=ORDER_GIVEN, raise_on_error=True, grouped=False):
        """specs: dicts with d (DeviceArray [n, m]), k and optional
        znorm_ddof (default -1 = already normalized), max_iter, init_mode,
        seed, init_labels, init_rows, convention, cal_select, s_precision,
        stop_on_collapse, labels_per_iter.  grouped: tcx_batch_run_grouped
        (one launch per phase over all jobs) instead of the worker pool.
        Returns one dict per job (labels, centroids, stats, status, message,
        worker, ms_queued, ms_run, ms_latency, and labels_per_iter when
        asked)."""
        jobs = (Job * len(specs))()
        keep = []
        outs = []
        for j, sp in zip(jobs, specs):
            d = sp["d"]
            n, m = d.shape
            k = sp["k"]
            j.d_x, j.n, j.m, j.k = d.ptr, n, m, k
            j.znorm_ddof = sp.get("znorm_ddof", -1)
            self.lib.tcx_fit_opts_default(ctypes.byref(j.opts))
            j.opts.max_iter = sp.get("max_iter", 10)
            j.opts.init_mode = sp.get("init_mode", 0)
            j.opts.seed = sp.get("seed", 0)
            j.opts.convention = sp.get("convention", CONV_CPU)
            j.opts.cal_select = sp.get("cal_select", 0)
            j.opts.s_precision = sp.get("s_precision", S_SPLIT_BF16)
            j.opts.stop_on_collapse = int(sp.get("stop_on_collapse", False))
            hist = None
            if sp.get("labels_per_iter"):
                hist = np.zeros((j.opts.max_iter, n), np.int32)
                keep.append(hist)
                j.opts.labels_per_iter = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
            for key, field in (("init_labels", "init_labels"), ("init_rows", "init_rows")):
                if sp.get(key) is not None:
                    a = np.ascontiguousarray(sp[key], np.int32)
                    keep.append(a)
                    setattr(j.opts, field, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
            lab = np.empty(n, np.int32)
            cent = np.empty((k, m), np.float32)
            keep += [lab, cent]
            j.h_labels, j.h_centroids = _ptr(lab), _ptr(cent)
            outs.append((lab, cent, hist))
        if grouped:
            rc = self.lib.tcx_batch_run_grouped(self.h, jobs, len(specs))
        else:
            rc = self.lib.tcx_batch_run(self.h, jobs, len(specs), order)
        if rc != OK and raise_on_error:
            check(rc)
        res = []
        for j, (lab, cent, hist) in zip(jobs, outs):
            r = {"labels": lab, "centroids": cent, "stats": j.stats.as_dict(),
                 "status": j.status, "message": j.message.decode(), "worker": j.worker,
                 "ms_queued": j.ms_queued, "ms_run": j.ms_run, "ms_latency": j.ms_latency}
            if hist is not None:
                r["labels_per_iter"] = hist[: j.stats.iterations]
            res.append(r)
        return res


class Server:
    """Serving loop: asynchronous fits and assignment queries against
    resident models (tcx_server_*)."""

    def __init__(self, device=0, fit_workers=4, query_workers=2, budget=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.tcx_server_create(device, fit_workers, query_workers, budget,
                                         ctypes.byref(h)))
        self.h = h
        _register(self, 0)

    def close(self):
        if self.h:
            check(self.lib.tcx_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_model(self, centroids):
        c = np.ascontiguousarray(centroids, np.float32)
        mid = ctypes.c_int()
        check(self.lib.tcx_server_add_model(self.h, _ptr(c), c.shape[0], c.shape[1],
                                            ctypes.byref(mid)))
        return mid.value

    def _submit(self, req, keep):
        check(self.lib.tcx_server_submit(self.h, ctypes.byref(req)))
        return {"req": req, "keep": keep}

    def submit_fit(self, X, k, znorm_ddof=0, publish=False, max_iter=10, seed=0, init_mode=0,
                   cal_select=0):
        X = np.ascontiguousarray(X, np.float32)
        n, m = X.shape
        r = Request()
        r.kind, r.h_x, r.n, r.m, r.znorm_ddof = REQ_FIT, _ptr(X), n, m, znorm_ddof
        r.k, r.publish_model = k, int(publish)
        self.lib.tcx_fit_opts_default(ctypes.byref(r.opts))
        r.opts.max_iter, r.opts.seed, r.opts.init_mode = max_iter, seed, init_mode
        r.opts.cal_select = cal_select
        lab = np.empty(n, np.int32)
        cent = np.empty((k, m), np.float32)
        r.h_labels, r.h_centroids = _ptr(lab), _ptr(cent)
        return self._submit(r, {"X": X, "labels": lab, "centroids": cent})

    def submit_assign(self, model, X, znorm_ddof=0):
        X = np.ascontiguousarray(X, np.float32)
        n, m = X.shape
        r = Request()
        r.kind, r.h_x, r.n, r.m, r.znorm_ddof, r.model = REQ_ASSIGN, _ptr(X), n, m, znorm_ddof, model
        lab = np.empty(n, np.int32)
        dist = np.empty(n, np.float32)
        sh = np.empty(n, np.int32)
        r.h_labels, r.h_dist, r.h_shift = _ptr(lab), _ptr(dist), _ptr(sh)
        return self._submit(r, {"X": X, "labels": lab, "dist": dist, "shift": sh})

    def wait(self, handle, raise_on_error=True):
        r = handle["req"]
        rc = self.lib.tcx_server_wait(self.h, ctypes.byref(r))
        if rc != OK and raise_on_error:
            check(rc)
        out = {key: v for key, v in handle["keep"].items() if key != "X"}
        out.update({"status": r.status, "message": r.message.decode(), "ms_queued": r.ms_queued,
                    "ms_run": r.ms_run, "ms_latency": r.ms_latency,
                    "model_out": r.model_out, "stats": r.stats.as_dict()})
        return out
