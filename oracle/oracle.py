"""ctypes bindings of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
module; it is the checker and the CPU baseline, never part of the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "sync_ref")

_lib = None


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)


class Geom(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("d", ctypes.c_int32), ("eps", ctypes.c_float),
                ("cs", ctypes.c_float), ("w", ctypes.c_int32), ("od", ctypes.c_int32),
                ("O", ctypes.c_int64), ("R", ctypes.c_int32), ("R_extra", ctypes.c_int32),
                ("R_half", ctypes.c_int32), ("eps_extra", ctypes.c_float),
                ("thr_eps", ctypes.c_float), ("thr_half", ctypes.c_float),
                ("thr_extra", ctypes.c_float)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64, i32, f32, f64 = ctypes.c_int64, ctypes.c_int32, ctypes.c_float, ctypes.c_double
        L.eo_geometry.argtypes = [i64, i32, f32, ctypes.POINTER(Geom)]
        L.eo_sqrt_threshold.argtypes = [f32]
        L.eo_sqrt_threshold.restype = f32
        L.eo_build_grid_f32.argtypes = [P, i64, i32, f32, P, P, P, P, P, P, P, P]
        L.eo_egg_iteration_f32.argtypes = [P, i64, i32, f32, P, P, P, P, P, P, P, P, P]
        L.eo_extra_check_f32.argtypes = [P, i64, i32, f32]
        L.eo_egg_solve_f32.argtypes = [P, i64, i32, f32, i32, P, P, P]
        L.eo_egg_solve_f64.argtypes = [P, i64, i32, f64, i32, P, P, P]
        L.eo_egg_labels_f32.argtypes = [P, i64, i32, f32, P]
        L.eo_sync_solve_f32.argtypes = [P, i64, i32, f32, f32, i32, P, P, P, P]
        L.eo_simple_labels_f32.argtypes = [P, i64, i32, f32, P]
        L.eo_sync_iteration_range_f32.argtypes = [P, i64, i32, f32, i64, i64, P, i32]
        L.eo_sync_iteration_range_f32.restype = f64
        L.eo_sync_labels_f32.argtypes = [P, i64, i32, f32, P]
        L.eo_rc_f64.argtypes = [P, i64, i32, f32]
        L.eo_rc_f64.restype = f64
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(X):
    return np.ascontiguousarray(X, dtype=np.float32)


def geometry(n: int, d: int, eps: float) -> Geom:
    g = Geom()
    if lib().eo_geometry(n, d, eps, ctypes.byref(g)) != 0:
        raise ValueError("bad geometry arguments")
    return g


def sqrt_threshold(e: float) -> float:
    return float(lib().eo_sqrt_threshold(e))


def build_grid(X, eps, order_in=None):
    X = _f32(X)
    n, d = X.shape
    g = geometry(n, d, eps)
    order = np.empty(n, np.int32)
    cell_of = np.empty(n, np.int32)
    cell_start = np.empty(n + 1, np.int32)
    cell_g = np.empty(max(n, 1) * d, np.int32)
    ob = np.empty(g.O, np.int32)
    oe = np.empty(g.O, np.int32)
    C = ctypes.c_int32()
    oi = None if order_in is None else np.ascontiguousarray(order_in, np.int32)
    rc = lib().eo_build_grid_f32(_p(X), n, d, eps, _p(oi), _p(order), _p(cell_of), _p(cell_start),
                                 _p(cell_g), _p(ob), _p(oe), ctypes.byref(C))
    if rc != 0:
        raise RuntimeError(f"eo_build_grid_f32 failed: {rc}")
    C = C.value
    return dict(order=order, cell_of=cell_of, cell_start=cell_start[: C + 1],
                cell_g=cell_g[: C * d].reshape(C, d), ob=ob, oe=oe, C=C, geom=g)


def egg_iteration(X, eps, order_in=None):
    X = _f32(X)
    n, d = X.shape
    X_next = np.empty_like(X)
    order = np.empty(n, np.int32)
    N = np.empty(n, np.int32)
    half = np.empty(n, np.int32)
    border = np.empty(n, np.int32)
    synced = np.empty(n, np.int32)
    fl = ctypes.c_int32()
    ff = ctypes.c_int32()
    oi = None if order_in is None else np.ascontiguousarray(order_in, np.int32)
    rc = lib().eo_egg_iteration_f32(_p(X), n, d, eps, _p(oi), _p(X_next), _p(order), _p(N),
                                    _p(half), _p(border), _p(synced), ctypes.byref(fl),
                                    ctypes.byref(ff))
    if rc != 0:
        raise RuntimeError(f"eo_egg_iteration_f32 failed: {rc}")
    return dict(X_next=X_next, order=order, N=N, half=half, border=border, synced=synced,
                flag_local=fl.value, flag_final=ff.value)


def extra_check(X, eps) -> int:
    X = _f32(X)
    return int(lib().eo_extra_check_f32(_p(X), X.shape[0], X.shape[1], eps))


def egg_solve(X, eps, max_iter=0, fp64=False):
    if fp64:
        X = np.ascontiguousarray(X, dtype=np.float64)
    else:
        X = _f32(X)
    n, d = X.shape
    labels = np.empty(n, np.int32)
    Xf = np.empty_like(X)
    it = ctypes.c_int32()
    fn = lib().eo_egg_solve_f64 if fp64 else lib().eo_egg_solve_f32
    st = fn(_p(X), n, d, eps, max_iter, _p(labels), _p(Xf), ctypes.byref(it))
    if st < 0:
        raise RuntimeError(f"egg solve failed: {st}")
    return dict(labels=labels, X=Xf, iterations=it.value, status=st)


def egg_labels(X, eps):
    X = _f32(X)
    labels = np.empty(X.shape[0], np.int32)
    lib().eo_egg_labels_f32(_p(X), X.shape[0], X.shape[1], eps, _p(labels))
    return labels


def sync_solve(X, eps, lam=np.float32(1 - 1e-3), max_iter=0):
    X = _f32(X)
    n, d = X.shape
    labels = np.empty(n, np.int32)
    Xf = np.empty_like(X)
    it = ctypes.c_int32()
    cap = max_iter if max_iter > 0 else 100000
    trace = np.zeros(cap, np.float32)
    st = lib().eo_sync_solve_f32(_p(X), n, d, eps, float(lam), max_iter, _p(labels), _p(Xf),
                                 ctypes.byref(it), _p(trace))
    if st < 0:
        raise RuntimeError(f"sync solve failed: {st}")
    return dict(labels=labels, X=Xf, iterations=it.value, r_trace=trace[: it.value], status=st)


def sync_iteration_range(X, eps, t0, t1, nthreads=0):
    X = _f32(X)
    n, d = X.shape
    Xn = X.copy()
    s = lib().eo_sync_iteration_range_f32(_p(X), n, d, eps, t0, t1, _p(Xn), nthreads)
    return Xn, s


def sync_labels(X, eps):
    X = _f32(X)
    labels = np.empty(X.shape[0], np.int32)
    lib().eo_sync_labels_f32(_p(X), X.shape[0], X.shape[1], eps, _p(labels))
    return labels


def simple_labels(X, eps):
    """Labels of the brute-force GPU SynC (simple_GPU_SynC.cu:70-102); returns (labels, k)."""
    X = _f32(X)
    labels = np.empty(X.shape[0], np.int32)
    k = lib().eo_simple_labels_f32(_p(X), X.shape[0], X.shape[1], eps, _p(labels))
    if k < 0:
        raise RuntimeError("simple labels failed")
    return labels, k


def simple_solve(X, eps, lam=np.float32(1 - 1e-3), max_iter=100):
    """GPU_DynamicalClustering (simple_GPU_SynC.cu:105-237): the SynC pass of
    SynC.cpp:16-56 (kernel_itr, :29-67, adds the same terms in the same order),
    r >= lam stop, at most max_iter passes (the reference: 100), then
    simple_labels.  The reference sums r with float atomics (order not fixed);
    here the FP32 sum of sync_solve."""
    r = sync_solve(X, eps, lam=lam, max_iter=max_iter)
    labels, k = simple_labels(r["X"], eps)
    r["labels"], r["n_clusters"] = labels, k
    return r


def rc(X, eps) -> float:
    X = _f32(X)
    return float(lib().eo_rc_f64(_p(X), X.shape[0], X.shape[1], eps))


def same_partition(a, b) -> bool:
    """Identical partitions up to relabelling (noise label -1 must match exactly)."""
    a = np.asarray(a).astype(np.int64)
    b = np.asarray(b).astype(np.int64)
    if a.shape != b.shape:
        return False
    if not np.array_equal(a < 0, b < 0):
        return False
    m = a >= 0
    if not m.any():
        return True
    pa, pb = a[m], b[m]
    pairs = np.unique(np.stack([pa, pb], 1), axis=0)
    return len(np.unique(pairs[:, 0])) == len(pairs) and len(np.unique(pairs[:, 1])) == len(pairs)
