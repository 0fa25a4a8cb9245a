"""The reference's experiment harness (run_experiment.py) on this solver.

Same axes and defaults as run_experiment.py:10-20 -- n, d, std, cl, eps,
three rounds per point, the six stage names -- the same synthetic data recipe
(run_experiment.py:274-286 -> python/data.py:24-43 -> python/generator.py:174-223:
cl Gaussian clusters in all d dimensions with sizes proportional to a random
integer in [5, 20], centres uniform in [-100, 100]^d, the remaining points
uniform in [-100, 100]^d, min-max normalised), seeded here so that a run can
be repeated, and loaders for the real data sets of python/data.py:46-97 that
read the same files with the same columns and delimiters.

    python -m egg_sync.experiments --vary n --values 2000 4000 8000 --rounds 3
    python -m egg_sync.experiments --vary real --data-dir <reference>/data

prints one JSON record per (value, round): running time (wall, around the
call, as run_experiment.py:322-325), per-iteration and per-stage times and
the device bytes held ("space").
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

PARAMS = {"n": 100_000, "d": 2, "std": 5.0, "cl": 5, "eps": 0.05, "rounds": 3}
SWEEPS = {
    "n": [2000, 4000, 8000, 16000, 32000, 64000, 128000, 256000, 512000, 1024000],
    "d": [2, 4, 6, 8, 10, 12, 14, 16, 18, 20],
    "std": [1.0, 5.0, 10.0, 15.0, 20.0],
    "cl": [2, 4, 8, 12, 16],
    "eps": [0.0125, 0.025, 0.050, 0.1, 0.2, 0.500],
}
STAGES = ["Allocating", "Build structure", "Update", "Extra check", "Clustering", "Free Memory"]


def min_max_normalize(X):
    """python/data.py:100-103 (a constant column maps to 0 here instead of NaN)."""
    X = np.asarray(X, dtype=np.float64)
    lo, hi = X.min(axis=0), X.max(axis=0)
    span = np.where(hi > lo, hi - lo, 1.0)
    return ((X - lo) / span).astype(np.float32)


def synthetic_gauss(n: int, d: int, cl: int, std: float, seed: int = 0, box: float = 100.0):
    """Gaussian clusters in all d dimensions plus uniform background, normalised."""
    rng = np.random.default_rng(seed)
    weights = rng.integers(5, 21, size=cl)
    sizes = (n * weights / weights.sum()).astype(np.int64)
    X = rng.uniform(-box, box, size=(n, d))
    start = 0
    for k in range(cl):
        centre = rng.uniform(-box, box, size=d)
        X[start:start + sizes[k]] = centre + rng.normal(0.0, std, size=(sizes[k], d))
        start += sizes[k]
    return min_max_normalize(X)


# Real data sets: python/data.py:46-97 (file name, loadtxt arguments).
DATASETS = {
    "iris": ("iris.data", dict(delimiter=",")),
    "pendigits": ("pendigits.tra", dict(delimiter=",")),
    "3D_spatial_network": ("3D_spatial_network.txt", dict(delimiter=",", usecols=(1, 2, 3))),
    "EEG_Eye_State": ("EEG Eye State.arff", dict(delimiter=",", usecols=tuple(range(14)))),
    "Letter": ("letter-recognition.data", dict(delimiter=",", usecols=tuple(range(1, 17)))),
    "CCPP": ("CCPP.data", dict(delimiter="\t", skiprows=1)),
    "banknote": ("data_banknote_authentication.txt", dict(delimiter=",", usecols=(0, 1, 2, 3))),
    "eb": ("eb.arff", dict(delimiter=",", usecols=(0, 1, 4))),
    "Skin_NonSkin": ("Skin_NonSkin.txt", dict(usecols=(0, 1, 2))),
    "Wilt": ("wilt.data", dict(delimiter=",", usecols=(1, 2, 3, 4, 5))),
    "Yeast": ("yeast.data", dict(usecols=(1, 2, 3, 4, 5, 6, 7, 8))),
}


# The real data sets of the reference's run_real (run_experiment.py:53-63,547-551).
REAL = ["banknote", "Yeast", "Wilt", "CCPP", "eb", "Skin_NonSkin", "3D_spatial_network",
        "EEG_Eye_State", "Letter"]


def load_dataset(name: str, data_dir: str):
    """A real data set from `data_dir`, min-max normalised, float32."""
    fname, kw = DATASETS[name]
    return min_max_normalize(np.loadtxt(os.path.join(data_dir, fname), **kw))


def run_real(data_dir: str, rounds: int = PARAMS["rounds"], eps: float = PARAMS["eps"],
             algorithm=None, names=None):
    """run_experiment.py:547-551: every real data set (those whose file is in
    `data_dir`; 3D_spatial_network is not shipped with the reference) at the
    default eps, `rounds` timed calls each."""
    if algorithm is None:
        from .algorithms import EGG_SynC as algorithm
    out = []
    for name in names or REAL:
        if not os.path.exists(os.path.join(data_dir, DATASETS[name][0])):
            continue
        X = load_dataset(name, data_dir)
        for r in range(rounds):
            secs, itr, stages, space = run_one(algorithm, X, eps)
            out.append({"dataset": name, "round": r, "n": int(X.shape[0]), "d": int(X.shape[1]),
                        "eps": float(eps), "running_time": secs, "iterations": int(len(itr)),
                        "stage_times": dict(zip(STAGES, [float(t) for t in stages])),
                        "space": int(space)})
    return out


def run_one(algorithm, X, eps: float):
    """One timed call, as run_experiment.py:322-325: (seconds, itr_times, stage_times, space)."""
    import torch
    D = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32))
    torch.cuda.synchronize()
    t0 = time.time()
    _, itr_times, stage_times, space = algorithm(D, eps)
    t1 = time.time()
    return t1 - t0, itr_times, stage_times, space


def run(vary: str = "n", values=None, rounds: int = PARAMS["rounds"], algorithm=None, seed=0,
        **overrides):
    """Records for every value of one axis and every round (run_experiment.py:294-331)."""
    if algorithm is None:
        from .algorithms import EGG_SynC as algorithm
    params = dict(PARAMS, **overrides)
    values = list(values) if values is not None else SWEEPS[vary]
    out = []
    for v in values:
        p = dict(params, **{vary: v})
        for r in range(rounds):
            X = synthetic_gauss(int(p["n"]), int(p["d"]), int(p["cl"]), float(p["std"]),
                                seed=seed + r)
            secs, itr, stages, space = run_one(algorithm, X, float(p["eps"]))
            out.append({"vary": vary, "value": v, "round": r, "n": int(p["n"]), "d": int(p["d"]),
                        "std": float(p["std"]), "cl": int(p["cl"]), "eps": float(p["eps"]),
                        "running_time": secs, "iterations": int(len(itr)),
                        "itr_times": [float(t) for t in itr],
                        "stage_times": dict(zip(STAGES, [float(t) for t in stages])),
                        "space": int(space)})
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--vary", default="n", choices=sorted(SWEEPS) + ["real"])
    ap.add_argument("--data-dir", default="data", help="directory of the real data files")
    ap.add_argument("--values", type=float, nargs="*")
    ap.add_argument("--rounds", type=int, default=PARAMS["rounds"])
    for k in ("n", "d", "cl"):
        ap.add_argument(f"--{k}", type=int, default=PARAMS[k])
    for k in ("std", "eps"):
        ap.add_argument(f"--{k}", type=float, default=PARAMS[k])
    a = ap.parse_args(argv)
    if a.vary == "real":
        for rec in run_real(a.data_dir, a.rounds, a.eps):
            print(json.dumps(rec), flush=True)
        return
    vals = None
    if a.values:
        vals = [int(v) if a.vary in ("n", "d", "cl") else v for v in a.values]
    fixed = {k: getattr(a, k) for k in ("n", "d", "std", "cl", "eps") if k != a.vary}
    for rec in run(a.vary, vals, a.rounds, **fixed):
        print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
