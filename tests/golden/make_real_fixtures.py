"""Writes tests/golden/real_oracle.npz: real data sets of the reference's run_real
(run_experiment.py:53-63) and pendigits (python/data.py:49-51, 17 columns), read
from the reference's data/ files with the loaders of egg_sync.experiments
(python/data.py:46-103: same files, columns and delimiters, min-max normalised),
each with the oracle's EGG-SynC solve at eps = 0.05 (run_experiment.py:12):
the normalised input, labels, iteration count and final positions.

    python tests/golden/make_real_fixtures.py /root/reference/data
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "egg-sync_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402
from egg_sync import experiments as E  # noqa: E402

NAMES = ["banknote", "Yeast", "Wilt", "CCPP", "pendigits"]
EPS = 0.05


def main(data_dir):
    O.build()
    out = {}
    for name in NAMES:
        X = E.load_dataset(name, data_dir)
        r = O.egg_solve(X, EPS)
        out[f"{name}_X"] = X
        out[f"{name}_labels"] = r["labels"]
        out[f"{name}_iterations"] = np.int32(r["iterations"])
        out[f"{name}_final"] = r["X"]
        print(name, X.shape, "iterations", r["iterations"], "clusters",
              int(r["labels"].max()) + 1, flush=True)
    np.savez_compressed(os.path.join(HERE, "real_oracle.npz"), eps=np.float32(EPS), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data")
