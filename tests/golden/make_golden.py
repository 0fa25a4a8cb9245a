"""Writes the golden fixtures of tests/golden/ from the reference's own CPU SynC.

Run in a container where /root/reference exists:
    bash oracle/build_ref.sh && python tests/golden/make_golden.py

It generates inputs with the project's seeded generators, runs
oracle/_ref/sync_ref (the reference's src/algorithms/SynC.cpp compiled from its
sources) on them, and stores inputs and outputs as .npz data:

  c0_sync.npz   config 0 (n=2000, d=2, eps=0.05): DynamicalClustering labels and
                iteration count, final positions and r trace of the replayed loop
  step_*.npz    one SynC pass (ComputeNeighborhood + UpdatePoint +
                ComputeLocationOrder) over every point: next positions,
                neighbourhood sizes, per-point r_p.  The EGG update computes the
                same map x -> x + mean_{N_eps} sin(x_j - x), so these pin it.
  digests.json  SHA-256 of small generator outputs (pins the generators)
  c0_fsync.npz  config 0 through the reference's FDynamicalClustering (FSynC: the same
                iteration with R-tree range queries, node capacity B = 100 as
                python/algorithms.py:98 passes): labels and iteration count

    python tests/golden/make_golden.py          # everything above
    python tests/golden/make_golden.py fsync    # c0_fsync.npz only
"""
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "egg-sync_amd"))
from egg_sync import data  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "sync_ref")
LAM = np.float32(1 - 1e-3)


def write_input(path, X, eps, lam=LAM):
    X = np.ascontiguousarray(X, np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<iiff", X.shape[0], X.shape[1], float(np.float32(eps)), float(lam)))
        f.write(X.tobytes())


def run_solve(X, eps):
    with tempfile.TemporaryDirectory() as td:
        inp, out = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(inp, X, eps)
        subprocess.run([REF_BIN, "solve", inp, out], check=True, stdout=subprocess.DEVNULL)
        buf = open(out, "rb").read()
    n, d, it_ref, it_rep = struct.unpack_from("<iiii", buf, 0)
    off = 16
    lab_ref = np.frombuffer(buf, np.int32, n, off); off += 4 * n
    lab_rep = np.frombuffer(buf, np.int32, n, off); off += 4 * n
    Xf = np.frombuffer(buf, np.float32, n * d, off).reshape(n, d); off += 4 * n * d
    tr = np.frombuffer(buf, np.float32, it_rep, off)
    return dict(labels_ref=lab_ref.copy(), labels_replay=lab_rep.copy(), iters_ref=it_ref,
                iters_replay=it_rep, X_final=Xf.copy(), r_trace=tr.copy())


def run_step(X, eps):
    with tempfile.TemporaryDirectory() as td:
        inp, out = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(inp, X, eps)
        subprocess.run([REF_BIN, "step", inp, out], check=True, stdout=subprocess.DEVNULL)
        buf = open(out, "rb").read()
    n, d = struct.unpack_from("<ii", buf, 0)
    off = 8
    Xn = np.frombuffer(buf, np.float32, n * d, off).reshape(n, d); off += 4 * n * d
    N = np.frombuffer(buf, np.int32, n, off); off += 4 * n
    rp = np.frombuffer(buf, np.float32, n, off)
    return dict(X_next=Xn.copy(), N=N.copy(), r_p=rp.copy())


def run_fsolve(X, eps, B):
    with tempfile.TemporaryDirectory() as td:
        inp, out = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(inp, X, eps)
        subprocess.run([REF_BIN, "fsolve", inp, out, str(B)], check=True,
                       stdout=subprocess.DEVNULL)
        buf = open(out, "rb").read()
    n, iters = struct.unpack_from("<ii", buf, 0)
    return np.frombuffer(buf, np.int32, n, 8).copy(), iters


def fsync():
    X0, eps0 = data.config0()
    labels, iters = run_fsolve(X0, eps0, 100)
    np.savez_compressed(os.path.join(HERE, "c0_fsync.npz"), X=X0, eps=np.float32(eps0),
                        lam=LAM, B=np.int32(100), labels=labels, iterations=np.int32(iters))
    print("c0_fsync: iterations", iters, "clusters", int(labels.max()) + 1)


def main():
    if not os.path.exists(REF_BIN):
        sys.exit("oracle/_ref/sync_ref missing: run oracle/build_ref.sh first")
    X0, eps0 = data.config0()
    s = run_solve(X0, eps0)
    assert s["iters_ref"] == s["iters_replay"], (s["iters_ref"], s["iters_replay"])
    assert np.array_equal(s["labels_ref"], s["labels_replay"])
    np.savez_compressed(os.path.join(HERE, "c0_sync.npz"), X=X0, eps=np.float32(eps0), lam=LAM,
                        labels=s["labels_ref"], iterations=np.int32(s["iters_ref"]),
                        X_final=s["X_final"], r_trace=s["r_trace"])
    print("c0_sync: iterations", s["iters_ref"], "clusters", int(s["labels_ref"].max()) + 1)

    steps = {
        "step_c0": (X0, eps0),
        "step_d3": (data.config1(n=5000)[0], np.float32(0.05)),
        "step_d5": (data.gaussian_blobs(3000, 5, 6, seed=11, center_lo=0.2, center_hi=0.8,
                                        sigma=0.05, noise=0.05), np.float32(0.15)),
        "step_d8": (data.config2(n=3000)[0], np.float32(0.25)),
    }
    # a mid-solve state of config 0 (after 3 passes) exercises tight clusters
    Xm = X0
    for _ in range(3):
        Xm = run_step(Xm, eps0)["X_next"]
    steps["step_c0_mid"] = (Xm, eps0)
    for name, (X, eps) in steps.items():
        r = run_step(X, eps)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), X=np.asarray(X, np.float32),
                            eps=np.float32(eps), X_next=r["X_next"], N=r["N"], r_p=r["r_p"])
        print(name, X.shape, "mean |N|", float(r["N"].mean()))

    digests = {
        "config0_n2000": data.digest(data.config0()[0]),
        "config1_n1000": data.digest(data.config1(n=1000)[0]),
        "config2_n1000": data.digest(data.config2(n=1000)[0]),
        "config3_first3": [data.digest(x) for x in data.config3_instances(count=3, n_hi=5000)],
    }
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(digests, f, indent=1)


if __name__ == "__main__":
    if not os.path.exists(REF_BIN):
        sys.exit("oracle/_ref/sync_ref missing: run oracle/build_ref.sh first")
    if sys.argv[1:] == ["fsync"]:
        fsync()
    else:
        main()
        fsync()
