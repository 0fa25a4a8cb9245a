"""Python API of the MI355X EGG-SynC, mirroring the reference's python/algorithms.py.

    EGG_SynC(D, eps)            python/algorithms.py:110-111 -> EGG_SynC.cu:1229
    SynC(D, eps)                python/algorithms.py:90-91   (r_c stop, lam = 1 - 1e-3)
    SynC_parallel(D, eps)       python/algorithms.py:94-95
    DynamicalClustering(D, eps, lam=1 - 1e-3, call=None)            :21-22
    DynamicalClustering_parallel(D, eps, lam=1 - 1e-3, call=None)   :25-26
    FSynC(D, eps, B=100)        python/algorithms.py:98-99
    FDynamicalClustering(D, eps, lam=1 - 1e-3, call=None, B=50)     :29-30
    simple_GPU_SynC(D, eps)     python/algorithms.py:106-107 (brute force, simple_GPU_SynC.cu)
    GPU_DynamicalClustering(D, eps, lam=1 - 1e-3, call=None)        :33-34

Each returns ``[labels, itr_times, stage_times, space]`` like the reference
(EGG_SynC.cu:1587-1600, SynC.cpp:309-325).  ``D`` is an [n, d] float32 tensor
with values in [0, 1] (use ``data.min_max_normalize``), on the CPU (copied in,
as the reference does) or already on the GPU (solved in place, on the current
stream).  The reference's CPU SynC variants run here on the GPU too: same
update, r_c stop rule and synCluster_v2 labels, different hardware.  FSynC is
SynC with an R-tree neighbour search (SynC.cpp:426-487); the neighbour sets
are the same, so it runs the same GPU solve (B, the tree's node capacity, is
checked and otherwise unused).

There is no CPU fallback: the calls raise if the native extension is missing.
"""
from __future__ import annotations

import importlib.machinery
import importlib.util
import os

import torch

from . import _native

_EXT = None


def _ext():
    global _EXT
    if _EXT is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_egg_sync_C.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: run `python -m egg_sync.build` first")
        _native.lib()  # the extension links libeggsync.so; fail with a clear message first
        loader = importlib.machinery.ExtensionFileLoader("_egg_sync_C", path)
        spec = importlib.util.spec_from_file_location("_egg_sync_C", path, loader=loader)
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
        _EXT = mod
    return _EXT


def _stream_of(D: torch.Tensor) -> int:
    return torch.cuda.current_stream(D.device).cuda_stream if D.is_cuda else 0


def EGG_SynC(D: torch.Tensor, eps: float):
    """EGG-SynC clustering of D (n x d in [0,1]) -> [labels(float32, -1 = noise), itr_times,
    stage_times, space]."""
    return _ext().EGG_SynC_wrapper(D, float(eps), _stream_of(D))


def DynamicalClustering(D: torch.Tensor, eps: float, lam: float = 1 - 1e-3, call=None):
    """SynC with the r_c >= lam stop rule and synCluster_v2 labels (int32)."""
    return _ext().DynamicalClustering_wrapper(D, float(eps), float(lam), _stream_of(D))


def DynamicalClustering_parallel(D: torch.Tensor, eps: float, lam: float = 1 - 1e-3, call=None):
    return _ext().DynamicalClustering_parallel_wrapper(D, float(eps), float(lam), _stream_of(D))


def FDynamicalClustering(D: torch.Tensor, eps: float, lam: float = 1 - 1e-3, call=None,
                         B: int = 50):
    """FSynC's solve (SynC.cpp:426-487): SynC with R-tree range queries; int32 labels."""
    return _ext().FDynamicalClustering_wrapper(D, float(eps), float(lam), int(B), _stream_of(D))


def FSynC(D: torch.Tensor, eps: float, B: int = 100):
    return FDynamicalClustering(D, eps, lam=1 - 1e-3, B=B)


def SynC(D: torch.Tensor, eps: float):
    return DynamicalClustering(D, eps, lam=1 - 1e-3)


def SynC_parallel(D: torch.Tensor, eps: float):
    return DynamicalClustering_parallel(D, eps, lam=1 - 1e-3)


def GPU_DynamicalClustering(D: torch.Tensor, eps: float, lam: float = 1 - 1e-3, call=None):
    """Brute-force SynC on the GPU: r_c >= lam stop with the reference's 100-pass cap,
    labels float32 (smallest eps-neighbour index, numbered 0..k-1; -1 = noise)."""
    return _ext().GPU_DynamicalClustering_wrapper(D, float(eps), float(lam), _stream_of(D))


def simple_GPU_SynC(D: torch.Tensor, eps: float):
    return GPU_DynamicalClustering(D, eps)
