"""Writes the full-size oracle fixtures, so GPU runs at full size are compared
with the CPU oracle without re-running it:

  c1_full_oracle.npz  config 1 (n=1,000,000, d=3, eps=0.02; ~7 min on 8 cores):
                      labels (int32, point-id order), iteration count, a seeded
                      sample of 20,000 final positions, column sums of all final
                      positions.
  c3_tail_oracle.npz  the slowest instance of the batched workload (dataset 27 of the
                      64-instance generator, n=165,583, d=2, eps=0.1; ~11,600
                      iterations, ~3 min): labels, iteration count, sampled final
                      positions.
  c1_fp64_oracle.npz  config 1 solved by the FP64 oracle: iteration count, labels, the
                      final positions at c1_full_oracle.npz's sample indices, and the
                      order parameter r_c of the final positions computed entirely in
                      FP64 (what the benchmark reports the GPU's r_c and positions against).
  c4_full_oracle.npz  config 4 (n=4,000,000, d=3, eps=0.01), run pass by pass: labels,
                      iteration count, sampled final positions, column sums, and per
                      pass the oracle's flags and CPU seconds.
  c2_full_oracle.npz  config 2 (n=250,000, d=8, eps=0.1; ~3 min): the same from the
                      FP32 oracle, plus the FP64 oracle's iteration count, labels
                      and sampled final positions, and the order parameter r_c of
                      the FP64 final positions computed entirely in FP64
                      (brute force over distinct positions with multiplicities).

    python tests/golden/make_oracle_fixtures.py [c1] [c1_fp64] [c2] [c3_tail] [c4]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "egg-sync_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402
from egg_sync import data  # noqa: E402


def rc_fp64(X, eps):
    """r = (1/n) sum_p (1/|N_p|) sum_{j in N_p} exp(-|x_j - x_p|), N_p = {j: |x_j - x_p| <= eps}
    (SynC.cpp:42-56), all in FP64, over distinct positions weighted by multiplicity."""
    U, inv, cnt = np.unique(np.asarray(X, np.float64), axis=0, return_inverse=True,
                            return_counts=True)
    r_u = np.empty(len(U))
    for a in range(0, len(U), 256):
        dist = np.sqrt(((U[a:a + 256, None, :] - U[None, :, :]) ** 2).sum(-1))
        w = (dist <= eps) * cnt[None, :]
        r_u[a:a + 256] = (w * np.exp(-dist)).sum(1) / w.sum(1)
    return float((r_u[inv.ravel()]).mean())


def c1():
    X, eps = data.config1()
    r = O.egg_solve(X, eps)
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(len(X), 20000, replace=False)).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "c1_full_oracle.npz"), labels=r["labels"],
                        iterations=np.int32(r["iterations"]), sample_idx=idx,
                        sample_X=r["X"][idx], X_sum=r["X"].astype(np.float64).sum(0))
    print("c1 iterations", r["iterations"], "clusters", int(r["labels"].max()) + 1, flush=True)


def c2():
    X, eps = data.config2()
    r = O.egg_solve(X, eps)
    print("c2 fp32 iterations", r["iterations"], flush=True)
    r64 = O.egg_solve(X, eps, fp64=True)
    print("c2 fp64 iterations", r64["iterations"], flush=True)
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(len(X), 5000, replace=False)).astype(np.int32)
    rc64 = rc_fp64(r64["X"], eps)
    np.savez_compressed(os.path.join(HERE, "c2_full_oracle.npz"), labels=r["labels"],
                        iterations=np.int32(r["iterations"]), sample_idx=idx,
                        sample_X=r["X"][idx], X_sum=r["X"].astype(np.float64).sum(0),
                        labels64=r64["labels"], iterations64=np.int32(r64["iterations"]),
                        sample_X64=r64["X"][idx], rc64=np.float64(rc64))
    print("c2 clusters", int(r["labels"].max()) + 1, "rc64", rc64, flush=True)


def c1_fp64():
    X, eps = data.config1()
    r = O.egg_solve(X, eps, fp64=True)
    idx = np.load(os.path.join(HERE, "c1_full_oracle.npz"))["sample_idx"]
    rc64 = rc_fp64(r["X"], eps)
    np.savez_compressed(os.path.join(HERE, "c1_fp64_oracle.npz"), labels=r["labels"],
                        iterations=np.int32(r["iterations"]), sample_idx=idx,
                        sample_X=r["X"][idx], rc64=np.float64(rc64))
    print("c1 fp64 iterations", r["iterations"], "clusters", int(r["labels"].max()) + 1,
          "rc64", rc64, flush=True)


def c4():
    import time
    X, eps = data.config4()
    cur, order = X, None
    flags, secs = [], []
    while True:
        t0 = time.perf_counter()
        it = O.egg_iteration(cur, eps, order)
        secs.append(time.perf_counter() - t0)
        flags.append((it["flag_local"], it["flag_final"]))
        cur, order = it["X_next"], it["order"]
        print(f"c4 pass {len(secs)}: {secs[-1]:.1f} s, flags {flags[-1]}", flush=True)
        if it["flag_final"]:
            break
    labels = O.egg_labels(cur, eps)
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(len(X), 40000, replace=False)).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "c4_full_oracle.npz"), labels=labels,
                        iterations=np.int32(len(secs)), sample_idx=idx, sample_X=cur[idx],
                        X_sum=cur.astype(np.float64).sum(0),
                        pass_flags=np.array(flags, np.int32),
                        pass_cpu_s=np.array(secs, np.float64))
    print("c4 iterations", len(secs), "clusters", int(labels.max()) + 1, flush=True)


def c3_tail():
    X = data.config3_instances(seed=3, count=64)[27]
    r = O.egg_solve(X, 0.1)
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(len(X), 5000, replace=False)).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "c3_tail_oracle.npz"), labels=r["labels"],
                        iterations=np.int32(r["iterations"]), sample_idx=idx,
                        sample_X=r["X"][idx])
    print("c3 tail iterations", r["iterations"], "clusters", int(r["labels"].max()) + 1,
          flush=True)


if __name__ == "__main__":
    O.build()
    todo = sys.argv[1:] or ["c1", "c2", "c3_tail"]
    for name in todo:
        {"c1": c1, "c1_fp64": c1_fp64, "c2": c2, "c3_tail": c3_tail, "c4": c4}[name]()
