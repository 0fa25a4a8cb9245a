"""Autotuner of the solver's launch shapes.

The knobs it searches change how work is launched, never what is computed:
the tile-queue kernels' workgroup count and size (`queue_blocks`,
`queue_threads`), iterations per host check (`iters_per_sync`), hipGraph
replay (`use_graph`) and, for batched solves, the number of streams.  Every
candidate's labels and final positions are compared bit for bit with the
first candidate's, and a candidate that differs is an error, not a result.

The search is coordinate descent over small candidate lists, each candidate
timed as the median of `repeats` solves after one untimed solve.  Winners are
cached in a JSON file (`EGG_SYNC_TUNE_CACHE`, default
~/.cache/egg_sync/tune.json) keyed by device name and a workload signature,
so later calls with a similar workload reuse them.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np
import torch

from . import solver

PLAIN_SPACE = {
    "queue_threads": [256, 128, 64],
    "queue_blocks": [0, 1024, 4096],
    "iters_per_sync": [4, 8, 16],
    "use_graph": [False, True],
}

BATCH_SPACE = {
    "streams": [4, 1, 2, 8, 16, 24, 32],
    "use_graph": [False, True],
    "iters_per_sync": [4, 2, 8],
    "queue_threads": [256, 128],
}


def cache_path() -> str:
    return os.environ.get("EGG_SYNC_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "egg_sync", "tune.json"))


def _device_key() -> str:
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|{p.multi_processor_count}"


def plain_signature(n: int, d: int, eps: float) -> str:
    return f"plain|n2^{int(round(math.log2(max(n, 1))))}|d{d}|eps{eps:.4g}"


def batch_signature(specs) -> str:
    pts = sum(int(s[1]) for s in specs)
    dims = sorted({int(s[2]) for s in specs})
    return (f"batch|k2^{int(round(math.log2(max(len(specs), 1))))}"
            f"|pts2^{int(round(math.log2(max(pts, 1))))}|d{','.join(map(str, dims))}")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(sig: str, best: dict, table):
    data = _load()
    data.setdefault(_device_key(), {})[sig] = {"best": best, "table": table}
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def cached(sig: str):
    """Best options stored for this device and signature, or None."""
    return _load().get(_device_key(), {}).get(sig, {}).get("best")


def _fingerprint(res) -> tuple:
    labels = res["labels"] if isinstance(res["labels"], list) else [res["labels"]]
    xs = res["X"] if isinstance(res["X"], list) else [res["X"]]
    return tuple(hash(l.tobytes()) for l in labels) + tuple(hash(x.tobytes()) for x in xs)


def _search(run, space: dict, repeats: int, log=None):
    cur = {k: v[0] for k, v in space.items()}
    table, ref = [], None
    timed = {}

    def measure(opts):
        key = json.dumps(opts, sort_keys=True)
        if key in timed:
            return timed[key]
        nonlocal ref
        res = run(opts)                       # untimed: captures graphs, sizes buffers
        fp = _fingerprint(res)
        if ref is None:
            ref = fp
        elif fp != ref:
            raise RuntimeError(f"launch options {opts} changed the result")
        ts = []
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run(opts)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        timed[key] = t
        table.append({"options": dict(opts), "seconds": t})
        if log:
            log(f"  {opts} -> {t * 1e3:.2f} ms")
        return t

    best_t = measure(cur)
    for k, values in space.items():
        for v in values:
            if v == cur[k]:
                continue
            cand = dict(cur, **{k: v})
            t = measure(cand)
            if t < best_t:
                best_t, cur = t, cand
    return cur, best_t, table


def tune_plain(X, eps: float, repeats: int = 3, space=None, store=True, log=None):
    """Tunes the single-solve launch shape on (X, eps).  Returns (best options, seconds, table)."""
    D = solver._dev_f32(X)

    def run(o):
        return solver.solve(D, eps, **o)

    best, t, table = _search(run, space or PLAIN_SPACE, repeats, log)
    if store:
        _store(plain_signature(D.shape[0], D.shape[1], eps), best, table)
    return best, t, table


def tune_batch(D: torch.Tensor, specs, repeats: int = 2, space=None, store=True, log=None):
    """Tunes the batched solve (streams, graph, iterations per check, queue
    workgroup size) on (D, specs).  Returns (best options, seconds, table)."""

    def run(o):
        return solver.solve_batch(D, specs, **o)

    best, t, table = _search(run, space or BATCH_SPACE, repeats, log)
    if store:
        _store(batch_signature(specs), best, table)
    return best, t, table
