"""Writes tests/golden/c1_full_oracle.npz: the CPU oracle's EGG-SynC solve of the
full config-1 workload (n=1,000,000, d=3, eps=0.02; about 7 minutes on 8 cores).

Stored: labels (int32, point-id order), iteration count, a seeded sample of
20,000 final positions and the column sums of all final positions, so a GPU
run at full size can be compared with the oracle without re-running it.
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


def main():
    X, eps = data.config1()
    r = O.egg_solve(X, eps)
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(len(X), 20000, replace=False)).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "c1_full_oracle.npz"), labels=r["labels"],
                        iterations=np.int32(r["iterations"]), sample_idx=idx,
                        sample_X=r["X"][idx], X_sum=r["X"].astype(np.float64).sum(0))
    print("iterations", r["iterations"], "clusters", int(r["labels"].max()) + 1)


if __name__ == "__main__":
    main()
