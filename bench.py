"""Benchmark: plain EGG-SynC solve of n=1,000,000 d=3 points at eps=0.02 on one MI355X.

A step is one complete solve (data tensor already in HBM -> labels in HBM).
value = point-updates/s = n * iterations * steps / timed seconds; the
time-to-clustering of one solve is ms_per_step.  The roofline object is for the
neighbour-update kernel (the dominant kernel): algorithmic bytes per launch
8*n*d + (12 + 8*d)*C (read x, write x', read each non-empty cell's start, end,
count and d sin / d cos totals once), divided by its mean launch duration from
HIP events recorded around every update launch on the solver's stream during
the timed steps.  The CPU baseline is the reference's own CPU SynC
(src/algorithms/SynC.cpp, compiled from source into oracle/_ref) running one
OpenMP pass over a bounded sample of target points of the same data.

The line also carries a "forms" object: the plain solve of the n=250k, d=8,
eps=0.1 workload; the learned-jump form on n=4M, d=3, eps=0.01 (trained
weights from egg_sync/weights, plain solve of the same data timed beside it,
labels compared); the batched form (64 datasets x 4 eps values = 256 solves
in one call).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
                    [--forms plain,plain_d8,learned,batched]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "egg-sync_amd"))

METRIC = "time-to-clustering (s), point-updates/s; plain vs learned form, same labels"
HBM_PEAK_GBS = 8000.0
VALU_PEAK = 256 * 2.4e9  # wave64 VALU instructions/s: 256 CUs x 2.4 GHz x 1 per CU per cycle
ITERS_PER_SYNC = 4
# Per update pass of the n=1,000,000 workload, from rocprofv3 counter runs of
# `bench.py --steps 1 --warmup 0 --no-cpu --forms plain` (mean over the solve's
# passes; the pass is the kernels listed under "kernel" below): HBM bytes
# (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads for a copy
# of known size) and VALU wave instructions (SQ_INSTS_VALU, which counts the
# pass's MFMA instructions too; those alone are UPDATE_MFMA_INSTS_N1M).
# Measured for this workload only, so reported only when n is 1,000,000.
UPDATE_PROFILE = "profiles/update_pass_config1_n1M_counters.md"
UPDATE_TRAFFIC_BYTES_N1M = 826.7e6
UPDATE_VALU_INSTS_N1M = 4.15e9
UPDATE_MFMA_INSTS_N1M = 2.75e8


def cpu_baseline(X, eps, target_seconds=12.0):
    """One SynC pass (SynC.cpp:352-361) over a sample of targets of the full data set."""
    import numpy as np
    n, d = X.shape
    threads = len(os.sched_getaffinity(0))
    ref = os.path.join(ROOT, "oracle", "_ref", "sync_ref")
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.bin")
        with open(inp, "wb") as f:
            f.write(np.array([n, d], np.int32).tobytes())
            f.write(np.array([eps, 1 - 1e-3], np.float32).tobytes())
            f.write(np.ascontiguousarray(X, np.float32).tobytes())

        def run_ref(t):
            out = subprocess.run([ref, "time", inp, "0", str(t), str(threads)], check=True,
                                 capture_output=True, text=True).stdout
            kv = dict(p.split("=") for p in out.split())
            return float(kv["seconds"])

        kind = "reference"
        try:
            if not os.path.exists(ref):
                raise FileNotFoundError(ref)
            probe_t = 4 * threads
            s = run_ref(probe_t)
            targets = int(min(n, max(probe_t, probe_t * target_seconds / max(s, 1e-3))))
            secs = run_ref(targets)
        except Exception:  # compiled reference unavailable: the oracle's restatement
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            kind = "port"
            probe_t = 4 * threads
            t0 = time.perf_counter()
            O.sync_iteration_range(X, eps, 0, probe_t, threads)
            s = time.perf_counter() - t0
            targets = int(min(n, max(probe_t, probe_t * target_seconds / max(s, 1e-3))))
            t0 = time.perf_counter()
            O.sync_iteration_range(X, eps, 0, targets, threads)
            secs = time.perf_counter() - t0
        base = {"value": targets / secs, "unit": "point-updates/s", "cores": threads, "kind": kind,
                "sample": f"one SynC pass (brute-force eps-neighbourhood, SynC.cpp:352-361) for "
                          f"the first {targets} of {n} points against all {n} points of the same "
                          f"data, {secs:.1f} s wall"}
        if kind == "reference":
            # the reference's tree-indexed CPU variant (FSynC, SynC.cpp:426-462): one
            # thread, R-tree with node capacity 100 (python/algorithms.py:98-99)
            try:
                def run_tree(t):
                    out = subprocess.run([ref, "ftime", inp, "0", str(t), "100"], check=True,
                                         capture_output=True, text=True).stdout
                    kv = dict(p.split("=") for p in out.split())
                    return float(kv["build_seconds"]), float(kv["seconds"])
                _, s = run_tree(8)
                tt = int(min(n, max(8, 8 * 0.5 * target_seconds / max(s, 1e-3))))
                build_s, tsecs = run_tree(tt)
                base["tree_index"] = {
                    "value": tt / tsecs, "unit": "point-updates/s", "cores": 1,
                    "kind": "reference",
                    "sample": f"FSynC's pass (R-tree range query + update, SynC.cpp:444-449) "
                              f"for the first {tt} of {n} points, {tsecs:.1f} s wall after a "
                              f"{build_s:.1f} s tree build over all {n} points"}
            except Exception as e:
                base["tree_index"] = {"error": f"{type(e).__name__}: {e}"}
    return base


WEIGHTS = os.path.join(ROOT, "egg-sync_amd", "egg_sync", "weights", "jump_d3.npz")


def learned_form(n=4_000_000):
    """Learned-jump solve vs plain solve of the same data (n x 3, eps=0.01).

    The learned solve is timed on its own (one plain pass, the jump, plain passes
    to EGG-SynC's stop rule) and its partition is compared with the plain one,
    next to how far two valid plain solves differing only in the order of their
    FP32 additions (tile_group 1 and 8) are from each other.  The verified form
    (learned solve, then the plain solve, whose labels are used when the
    partitions differ) costs both."""
    import torch
    from egg_sync import data, learned, solver
    from egg_sync.metrics import disagreement, same_partition
    if not os.path.exists(WEIGHTS):
        return {"skipped": f"no trained weights at {os.path.relpath(WEIGHTS, ROOT)}"}
    net = learned.JumpMLP.from_file(WEIGHTS)
    print("learned form: warm-up", file=sys.stderr, flush=True)
    X, eps = data.config4(n=n)
    D = torch.from_numpy(X).cuda()
    jump = dict(t0=1, alpha=0.5, clip=0.5)
    learned.solve_learned(D, eps, net, **jump)  # warm-up
    solver.solve(D, eps, return_positions=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p = solver.solve(D, eps, return_positions=False)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    q = solver.solve(D, eps, return_positions=False, tile_group=8)
    r = learned.solve_learned(D, eps, net, **jump)
    same = bool(same_partition(r["labels"], p["labels"]))
    return {"workload": f"n={n}, d=3, eps={eps}",
            "jump": "t0=1 plain pass, jump by 0.5 eps x the prediction clamped to +-0.5",
            "plain_s": t_plain, "plain_iterations": p["stats"]["iterations"],
            "plain_point_updates_per_s": n * p["stats"]["iterations"] / t_plain,
            "learned_s": r["seconds"], "learned_passes": r["iterations"],
            "speedup_unverified": t_plain / r["seconds"],
            "disagreement_with_plain": disagreement(r["labels"], p["labels"]),
            "plain_order_sensitivity": disagreement(q["labels"], p["labels"]),
            "same_partition": same,
            "verified_s": r["seconds"] + t_plain,
            "result_used": "learned" if same else "plain (fallback)",
            "meets_identical_labels_target": same and t_plain / r["seconds"] >= 3.0}


def plain_d8_form():
    """The plain solve on the n=250,000, d=8, eps=0.1 workload (sparse 8-d grid)."""
    import torch
    from egg_sync import data, solver
    X, eps = data.config2()
    D = torch.from_numpy(X).cuda()
    solver.solve(D, eps, return_positions=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = solver.solve(D, eps, return_positions=False)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    it = r["stats"]["iterations"]
    return {"workload": f"n={len(X)}, d=8, eps={eps}", "time_to_clustering_s": t,
            "iterations": it, "point_updates_per_s": len(X) * it / t,
            "clusters": r["stats"]["n_clusters"]}


def batched_form(count=64, streams=24):
    """256 solves (64 datasets x 4 eps) in one batched call vs one call each.  The
    batch deals its instances over `streams` HIP streams (24: the fastest of 1, 4,
    8, 12, 16, 24, 32, 48 and 64 measured on this workload)."""
    import torch
    from egg_sync import data, solver
    inst = data.config3_instances(seed=3, count=count)
    D, specs = solver.pack_instances(inst, [list(data.CONFIG3_EPS)] * len(inst))
    print(f"batched form: {len(specs)} solves, warm-up", file=sys.stderr, flush=True)
    solver.solve_batch(D, specs, streams=streams, return_positions=False)
    torch.cuda.synchronize()
    print("batched form: timed batch", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    r = solver.solve_batch(D, specs, streams=streams, return_positions=False)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    updates = sum(int(s[1]) * res["iterations"] for s, res in zip(specs, r["results"]))
    print("batched form: one call per solve", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    singles_same = True
    for k, (o, nk, dk, e) in enumerate(specs):
        x = D[o:o + nk * dk].view(nk, dk)
        q = solver.solve(x, e, return_positions=False)
        singles_same &= bool((q["labels"] == r["labels"][k]).all())
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    return {"workload": f"{len(specs)} solves: {count} datasets (n log-uniform in [1e3,2e5], "
                        "d in {2,3,5,8}) x eps {0.01,0.02,0.05,0.1}",
            "points": int(sum(int(s[1]) for s in specs)), "streams": streams,
            "batch_s": t_batch,
            "point_updates_per_s": updates / t_batch, "one_call_each_s": t_single,
            "speedup_vs_one_call_each": t_single / t_batch,
            "labels_equal_to_single_calls": singles_same,
            "max_iterations": max(res["iterations"] for res in r["results"])}


def copy_bandwidth(nbytes=1 << 30, reps=5):
    """Achievable HBM bandwidth: read + write bytes of a device-to-device copy / time."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (a.elapsed_time(b) / 1000.0) / 1e9
    del src, dst
    return gbs


def resident_solver(D, eps, iters_per_sync):
    """Device-resident solve through the C ABI (egg_sync_solve_device): input and
    labels stay in HBM, only the loop state (a few dozen bytes per batch of
    iterations) crosses to the host."""
    import ctypes
    import numpy as np
    import torch
    from egg_sync import _native as nat
    n, d = D.shape
    labels = torch.empty(n, dtype=torch.int32, device=D.device)
    itr = np.zeros(4096, np.float32)
    cells = np.zeros(4096, np.int32)

    def run(timing):
        opt = nat.default_options(timing=int(timing), iters_per_sync=iters_per_sync)
        st = nat.Stats()
        st.itr_ms = itr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        st.itr_ms_cap = len(itr)
        st.trace_cap = len(cells)
        st.cells_trace = cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        nat.check(nat.lib().egg_sync_solve_device(nat.ptr(D), n, d, float(eps),
                                                  ctypes.byref(opt), nat.ptr(labels), None,
                                                  ctypes.byref(st), stream),
                  "egg_sync_solve_device")
        it = st.iterations
        return {"iterations": it, "n_clusters": st.n_clusters,
                "update_kernel_ms": st.update_kernel_ms, "update_launches": st.update_launches,
                "stage_ms": list(st.stage_ms), "itr_ms": itr[:it].copy() if timing else None,
                "cells_trace": cells[:min(it, len(cells))].copy()}
    return run, labels


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--iters-per-sync", type=int, default=ITERS_PER_SYNC)
    ap.add_argument("--forms", default="plain,plain_d8,learned,batched",
                    help="comma list; plain is always measured")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a kernel of known traffic "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the solve runs on one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU solve would split the outer grid into slabs along dimension 0 "
              "with a halo of R outer cells; it is not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from egg_sync import algorithms, data

    if args.copy_calibration:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    X, eps = data.config1(n=args.n)
    n, d = X.shape
    D = torch.from_numpy(X).cuda()
    torch.cuda.synchronize()
    run, labels_dev = resident_solver(D, eps, args.iters_per_sync)

    for _ in range(args.warmup):
        run(True)
    torch.cuda.synchronize()
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        results.append(run(True))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    labels0 = labels_dev.cpu().numpy()

    iters = [r["iterations"] for r in results]
    updates = sum(n * it for it in iters)
    value = updates / wall
    # roofline of the update pass over the timed launches (HIP events on the
    # solver's stream around every pass)
    upd_ms = sum(r["update_kernel_ms"] for r in results)
    launches = sum(r["update_launches"] for r in results)
    cells = np.concatenate([r["cells_trace"] for r in results]).astype(np.float64)
    bytes_per_launch = float(np.mean(8.0 * n * d + (12.0 + 8.0 * d) * cells))
    avg_launch_s = upd_ms / 1000.0 / max(launches, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    st = results[-1]

    # the reference's own call, host tensor in and host labels out
    # (run_experiment.py:321-325 times EGG_SynC(D_cpu, eps) this way)
    Xc = torch.from_numpy(X)
    algorithms.EGG_SynC(Xc, eps)
    host_s = []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        lab_h, _, _, _ = algorithms.EGG_SynC(Xc, eps)
        host_s.append(time.perf_counter() - t1)
    same_host = bool(np.array_equal(lab_h.numpy() >= 0, labels0 >= 0))
    full_size = n == 1_000_000
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 20-blob Gaussian mixture (Zipf 1.1 sizes, per-axis sigma "
                "U[0.005,0.02]) + 2% uniform noise, min-max normalised to [0,1]^3",
        "config": {"workload": f"plain EGG-SynC solve, n={n}, d={d}, eps={eps}; a step is one "
                               "solve from the input in HBM to the labels in HBM "
                               "(egg_sync_solve_device)",
                   "form": "plain", "n": n, "d": d, "eps": eps, "iterations": iters[0],
                   "iters_per_sync": args.iters_per_sync,
                   "time_to_clustering_s": wall / args.steps,
                   "time_to_clustering_host_s": float(np.median(host_s)),
                   "host_call": "EGG_SynC(D_cpu, eps) of the drop-in API: host tensor in, host "
                                "labels out, 12 MB copied each way (run_experiment.py:321-325)",
                   "host_point_updates_per_s": n * iters[0] / float(np.median(host_s)),
                   "host_noise_mask_equal": same_host,
                   "clusters": st["n_clusters"],
                   "stage_ms_last_step": [round(x, 3) for x in st["stage_ms"]],
                   "iteration_ms_last_step": [round(float(x), 3) for x in st["itr_ms"]],
                   "cells_per_iteration_last_step": [int(c) for c in st["cells_trace"]]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": UPDATE_TRAFFIC_BYTES_N1M if full_size else None,
                     "traffic_source": UPDATE_PROFILE if full_size else None,
                     "kernel": "update pass: k_update_tile<3> + k_update_heavy<3> + "
                               "k_update_finish<3> + k_run_boxes<3> + k_super_boxes<3> + "
                               "k_update_pairs<3> + k_scatter_runs<3>",
                     "launches": launches,
                     "avg_launch_us": avg_launch_s * 1e6,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
    }
    if full_size:
        # the order parameter r_c of the final state (computed in FP64 on the
        # GPU, SynC.cpp:42-56) and the final positions, against a run of the
        # oracle entirely in FP64 (committed fixture, not run here)
        from egg_sync import solver
        from egg_sync.metrics import same_partition
        a = solver.solve(D, eps, compute_rc=True)
        g = np.load(os.path.join(ROOT, "tests", "golden", "c1_fp64_oracle.npz"))
        idx = g["sample_idx"]
        line["fp64_check"] = {
            "rc": float(a["stats"]["rc"]), "rc_fp64_oracle": float(g["rc64"]),
            "rc_abs_err": abs(float(a["stats"]["rc"]) - float(g["rc64"])),
            "max_abs_pos_err": float(np.abs(a["X"][idx].astype(np.float64)
                                            - g["sample_X"]).max()),
            "positions_compared": int(len(idx)),
            "iterations": int(a["stats"]["iterations"]),
            "iterations_fp64_oracle": int(g["iterations"]),
            "same_partition_as_fp64_oracle": bool(same_partition(a["labels"], g["labels"])),
            "source": "tests/golden/c1_fp64_oracle.npz (tests/golden/make_oracle_fixtures.py "
                      "c1_fp64: the oracle's solve with every operation in FP64)"}
        # the pass is bound by VALU instruction issue, not by HBM bytes: wave
        # instructions per pass (rocprofv3 SQ_INSTS_VALU over the same solve)
        # over the pass time measured here, against 256 CUs x 2.4 GHz x one
        # wave64 VALU instruction per CU per cycle
        valu = UPDATE_VALU_INSTS_N1M / avg_launch_s
        line["roofline"]["valu"] = {"achieved": valu, "peak": VALU_PEAK, "unit": "wave-instr/s",
                                    "frac": valu / VALU_PEAK, "instructions_per_launch":
                                    UPDATE_VALU_INSTS_N1M, "mfma_instructions_per_launch":
                                    UPDATE_MFMA_INSTS_N1M, "source": UPDATE_PROFILE}
    try:
        cp = copy_bandwidth()
        line["roofline"]["copy_gbs_measured"] = cp
        line["roofline"]["frac_of_measured_copy"] = achieved / cp
    except Exception as e:
        line["roofline"]["copy_gbs_measured"] = None
        print(f"copy bandwidth: {e}", file=sys.stderr)
    forms = {}
    for name, fn in (("plain_d8", plain_d8_form), ("learned", learned_form),
                     ("batched", batched_form)):
        if name in args.forms.split(","):
            try:
                forms[name] = fn()
            except Exception as e:  # report, keep the plain line
                forms[name] = {"error": f"{type(e).__name__}: {e}"}
            print(f"form {name}: {forms[name]}", file=sys.stderr, flush=True)
    if forms:
        line["forms"] = forms
    if not args.no_cpu:
        print("cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(X, eps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
