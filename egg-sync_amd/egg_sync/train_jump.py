"""Trains the learned-jump network on trajectories of the plain solver.

    python -m egg_sync.train_jump --runs 200 --n 100000 --eps 0.01 0.02 \\
        --out egg-sync_amd/egg_sync/weights/jump_d3.npz

Each run draws a dataset from the n=1M config generator (seeded, scaled to
`--n` points), solves it with the plain GPU solver, and keeps, for every
point, the features after `--t0` plain passes and the remaining displacement
to the converged position (in units of eps).  The network is trained with the
library's HIP forward/backward/Adam kernels on a random 90% of the runs and
scored on the other 10% (mean squared jump error next to that of predicting
no jump).  Weights and the training record go to one .npz file.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=200)
    ap.add_argument("--n", type=int, default=100_000)
    ap.add_argument("--eps", type=float, nargs="+", default=[0.01, 0.02])
    ap.add_argument("--t0", type=int, default=1)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    from . import data, learned

    t_start = time.perf_counter()
    Fs, Ts = [], []
    for r in range(a.runs):
        X, _ = data.config1(seed=1000 + r, n=a.n)
        eps = a.eps[r % len(a.eps)]
        F, T = learned.trajectory_samples(X, eps, t0=a.t0)
        Fs.append(F)
        Ts.append(T)
        if r % 20 == 0 or r == a.runs - 1:
            print(f"trajectories {r + 1}/{a.runs}  {time.perf_counter() - t_start:.1f} s",
                  flush=True)
    rng = np.random.default_rng(a.seed)
    runs = rng.permutation(a.runs)
    n_val = max(1, a.runs // 10)
    val, tr = runs[:n_val], runs[n_val:]
    F = torch.cat([Fs[i] for i in tr])
    T = torch.cat([Ts[i] for i in tr])
    Fv = torch.cat([Fs[i] for i in val])
    Tv = torch.cat([Ts[i] for i in val])
    del Fs, Ts
    net = learned.JumpMLP(3, seed=a.seed)
    t_fit = time.perf_counter()
    losses = net.fit(F, T, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                     log=lambda m: print(m, flush=True))
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t_fit
    pred = torch.cat([net.forward(Fv[i:i + 1 << 20]) for i in range(0, Fv.shape[0], 1 << 20)])
    val_mse = float(((pred - Tv) ** 2).sum(1).mean())
    base_mse = float((Tv ** 2).sum(1).mean())
    rec = dict(runs=a.runs, n=a.n, eps=a.eps, t0=a.t0, epochs=a.epochs, batch=a.batch, lr=a.lr,
               seed=a.seed, samples=int(F.shape[0]), train_loss=losses, val_mse=val_mse,
               val_mse_no_jump=base_mse, fit_seconds=fit_s,
               total_seconds=time.perf_counter() - t_start)
    np.savez(a.out, params=net.state(), n_in=net.n_in, hidden=net.h, n_out=net.n_out,
             record=json.dumps(rec))
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
