"""ctypes binding of the C ABI in include/egg_sync.h (libeggsync.so).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libeggsync.so")

OK, CAPPED = 0, 1
E_ARG, E_RANGE, E_DIM, E_KEY, E_ALLOC, E_HIP, E_NODEV, E_TOO_BIG, E_INTERNAL = (
    -1, -2, -3, -4, -5, -6, -7, -8, -9)
STOP_EGG, STOP_RC = 0, 1
LABELS_CELLS, LABELS_SYNC = 0, 1
MAX_D = 20

EXPORTED = [
    "egg_sync_default_options", "egg_sync_solve", "egg_sync_solve_ex", "egg_sync_solve_device",
    "egg_sync_build_grid_device", "egg_sync_iteration_device", "egg_sync_rc_device",
    "egg_sync_geometry", "egg_sync_release", "egg_sync_error_string",
    "egg_sync_features_device", "egg_sync_jump_device", "egg_mlp_param_count",
    "egg_mlp_workspace_floats", "egg_mlp_forward_device", "egg_mlp_train_step_device",
    "egg_sync_solve_batch_device", "egg_sync_simple_solve_device", "egg_sync_simple_solve",
    "egg_sync_simple_labels_device", "egg_sync_solve_device_noalloc",
]


class Options(ctypes.Structure):
    _fields_ = [("stop_rule", ctypes.c_int32), ("label_rule", ctypes.c_int32),
                ("lam", ctypes.c_float), ("max_iter", ctypes.c_int32),
                ("compute_rc", ctypes.c_int32), ("timing", ctypes.c_int32),
                ("iters_per_sync", ctypes.c_int32), ("validate", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("queue_blocks", ctypes.c_int32),
                ("queue_threads", ctypes.c_int32), ("streams", ctypes.c_int32),
                ("tile_group", ctypes.c_int32), ("pair_groups", ctypes.c_int32)]


class Instance(ctypes.Structure):
    _fields_ = [("data_offset", ctypes.c_int64), ("n", ctypes.c_int64), ("d", ctypes.c_int32),
                ("eps", ctypes.c_float)]


class InstanceResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("status", ctypes.c_int32),
                ("n_clusters", ctypes.c_int32), ("n_cells", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("status", ctypes.c_int32),
                ("n_clusters", ctypes.c_int32), ("n_cells", ctypes.c_int32),
                ("rc", ctypes.c_double), ("stage_ms", ctypes.c_double * 6),
                ("total_ms", ctypes.c_double), ("update_kernel_ms", ctypes.c_double),
                ("update_launches", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("workspace_bytes", ctypes.c_int64), ("itr_ms", ctypes.POINTER(ctypes.c_float)),
                ("itr_ms_cap", ctypes.c_int32), ("trace_cap", ctypes.c_int32),
                ("cells_trace", ctypes.POINTER(ctypes.c_int32)),
                ("rc_trace", ctypes.POINTER(ctypes.c_double))]


class EggSyncError(RuntimeError):
    def __init__(self, code: int, where: str = ""):
        self.code = code
        super().__init__(f"{where}: {error_string(code)} ({code})" if where else error_string(code))


_lib = None


def lib():
    """The loaded libeggsync.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m egg_sync.build` "
                              "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        P, i64, i32, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float
        L.egg_sync_default_options.argtypes = [ctypes.POINTER(Options)]
        L.egg_sync_solve.argtypes = [P, i64, i32, f32, i32, P, ctypes.POINTER(Stats)]
        L.egg_sync_solve_ex.argtypes = [P, i64, i32, f32, ctypes.POINTER(Options), P, P,
                                        ctypes.POINTER(Stats)]
        L.egg_sync_solve_device.argtypes = [P, i64, i32, f32, ctypes.POINTER(Options), P, P,
                                            ctypes.POINTER(Stats), P]
        L.egg_sync_solve_device_noalloc.argtypes = [P, i64, i32, f32, ctypes.POINTER(Options),
                                                    P, P, ctypes.POINTER(Stats), P]
        L.egg_sync_build_grid_device.argtypes = [P, i64, i32, f32, P, P, P, P, P, P, P,
                                                 ctypes.POINTER(ctypes.c_int32), P]
        L.egg_sync_iteration_device.argtypes = [P, i64, i32, f32, P, P, P, P, P, P, P,
                                                ctypes.POINTER(ctypes.c_int32),
                                                ctypes.POINTER(ctypes.c_int32), P]
        L.egg_sync_rc_device.argtypes = [P, i64, i32, f32, ctypes.POINTER(ctypes.c_double), P]
        L.egg_sync_geometry.argtypes = [i64, i32, f32, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_float)]
        L.egg_sync_release.argtypes = []
        L.egg_sync_features_device.argtypes = [P, P, i64, i32, f32, P, P]
        L.egg_sync_jump_device.argtypes = [P, P, i64, i32, f32, f32, P]
        L.egg_mlp_param_count.argtypes = [i32, i32, i32, ctypes.POINTER(ctypes.c_int64)]
        L.egg_mlp_workspace_floats.argtypes = [i32, i32, i32, i64]
        L.egg_mlp_workspace_floats.restype = ctypes.c_size_t
        L.egg_mlp_forward_device.argtypes = [P, i32, i32, i32, P, i64, P, P]
        L.egg_mlp_train_step_device.argtypes = [P, P, P, i32, i32, i32, P, P, i64, f32, i32, P,
                                                i64, P, ctypes.POINTER(ctypes.c_double), P]
        L.egg_sync_solve_batch_device.argtypes = [P, ctypes.POINTER(Instance), i32,
                                                  ctypes.POINTER(Options), P, P,
                                                  ctypes.POINTER(InstanceResult),
                                                  ctypes.POINTER(Stats), P]
        L.egg_sync_simple_solve_device.argtypes = [P, i64, i32, f32, f32, i32, P, P,
                                                   ctypes.POINTER(Stats), P]
        L.egg_sync_simple_solve.argtypes = [P, i64, i32, f32, f32, i32, P, P,
                                            ctypes.POINTER(Stats)]
        L.egg_sync_simple_labels_device.argtypes = [P, i64, i32, f32, P,
                                                    ctypes.POINTER(ctypes.c_int32), P]
        L.egg_sync_error_string.argtypes = [ctypes.c_int]
        L.egg_sync_error_string.restype = ctypes.c_char_p
        _lib = L
    return _lib


def error_string(code: int) -> str:
    return lib().egg_sync_error_string(code).decode()


def check(code: int, where: str = "") -> int:
    if code < 0:
        raise EggSyncError(code, where)
    return code


def default_options(**kw) -> Options:
    o = Options()
    check(lib().egg_sync_default_options(ctypes.byref(o)))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


def geometry(n: int, d: int, eps: float) -> dict:
    out = (ctypes.c_int64 * 5)()
    cs = ctypes.c_float()
    check(lib().egg_sync_geometry(n, d, eps, out, ctypes.byref(cs)), "geometry")
    return dict(w=out[0], od=out[1], O=out[2], R=out[3], key_bits=out[4], cs=cs.value)


def ptr(t) -> ctypes.c_void_p:
    """Device or host address of a torch tensor / numpy array (None -> NULL)."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(t.data_ptr())
