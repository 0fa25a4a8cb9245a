"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

Bar: integer results (sorted order, cells, ranges, neighbour counts, flags,
iteration counts) bit-exact; labels identical partitions; FP32 positions
within 1e-6 after one pass and 1e-5 after a full solve.  Full-size runs are
checked through size-independent properties (determinism, bit-exact pass on
the converged state, labels equal to the cell rule on the final positions).
"""
import sys

import numpy as np
import pytest
import torch

from conftest import golden

pytestmark = pytest.mark.gpu

f32 = np.float32


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def S():
    assert torch.cuda.is_available()
    from egg_sync import solver
    return solver


def _cases_grid():
    from egg_sync import data
    rng = np.random.default_rng(7)
    cases = [("c0", *data.config0()),
             ("c1_20k", data.config1(n=20000)[0], 0.02),
             ("c2_5k", data.config2(n=5000)[0], 0.1)]
    # d >= 9: keys of 3 and 4 words (d=20 eps=0.05: 192 sort bits, eps=0.02: 216),
    # and without split bits (d=19 eps=0.002: 247 cell bits + 9 hash bits)
    for d, eps in [(1, 0.01), (4, 0.1), (5, 0.15), (6, 0.2), (7, 0.3), (8, 0.01), (9, 0.2),
                   (12, 0.3), (17, 0.4), (20, 0.05), (20, 0.02), (19, 0.002)]:
        X = rng.random((3000, d)).astype(f32)
        X[:3] = 1.0
        X[3:6] = 0.0
        cases.append((f"rand_d{d}_eps{eps}", X, eps))
    for d, eps in [(12, 0.3), (20, 0.3)]:
        X = data.gaussian_blobs(3000, d, 5, seed=d, center_lo=0.2, center_hi=0.8, sigma=0.03,
                                noise=0.05)
        X[:40] = X[40]   # a run of equal positions
        cases.append((f"blobs_d{d}", X, eps))
    return cases


@pytest.mark.parametrize("case", _cases_grid(), ids=lambda c: c[0])
def test_grid_bit_exact(S, O, case):
    name, X, eps = case
    _say("grid", name, X.shape, eps)
    rng = np.random.default_rng(3)
    for order_in in (None, rng.permutation(len(X)).astype(np.int32)):
        a = S.build_grid(X, eps, order_in)
        b = O.build_grid(X, eps, order_in)
        assert a["C"] == b["C"]
        for k in ("order", "cell_of", "cell_start", "cell_g", "ob", "oe"):
            assert np.array_equal(a[k], b[k]), k


def _cases_iter():
    from egg_sync import data
    cases = [(nm, golden(nm)["X"], float(golden(nm)["eps"]))
             for nm in ("step_c0", "step_c0_mid", "step_d3", "step_d5", "step_d8")]
    cases.append(("c1_30k", data.config1(n=30000)[0], 0.02))
    cases.append(("c2_4k", data.config2(n=4000)[0], 0.1))
    X0, eps0 = data.config0()
    cases.append(("c0_tight", X0, 0.01))
    for d, eps in [(12, 0.3), (20, 0.3)]:
        X = data.gaussian_blobs(3000, d, 5, seed=d, center_lo=0.2, center_hi=0.8, sigma=0.03,
                                noise=0.05)
        cases.append((f"blobs_d{d}", X, eps))
    return cases


@pytest.mark.parametrize("case", _cases_iter(), ids=lambda c: c[0])
def test_iteration_bit_exact(S, O, case):
    name, X, eps = case
    _say("iteration", name, X.shape, eps)
    a = S.iteration(X, eps)
    b = O.egg_iteration(X, eps)
    for k in ("order", "N", "half", "border", "synced"):
        assert np.array_equal(a[k], b[k]), k
    assert (a["flag_local"], a["flag_final"]) == (b["flag_local"], b["flag_final"])
    err = float(np.abs(a["X_next"] - b["X_next"]).max())
    assert err <= 1e-6, err


def _advance(O, X, eps, k):
    for _ in range(k):
        X = O.egg_iteration(X, eps)["X_next"]
    return X


def test_iteration_bit_exact_late_states(S, O):
    """States after several passes: collapsing clusters, border points, extra check."""
    from egg_sync import data
    X, eps = data.config0()
    for k in (2, 5, 8):
        Xk = _advance(O, X, eps, k)
        _say("late state", k)
        a = S.iteration(Xk, eps)
        b = O.egg_iteration(Xk, eps)
        for key in ("order", "N", "half", "border", "synced"):
            assert np.array_equal(a[key], b[key]), (k, key)
        assert (a["flag_local"], a["flag_final"]) == (b["flag_local"], b["flag_final"])
        assert np.abs(a["X_next"] - b["X_next"]).max() <= 1e-6


def test_iteration_bit_exact_many_chunk_cells(S, O):
    """Cells of tens of thousands of points: a cell's summary is the combine of
    up to ~80 partial sums (more than one per lane of a wavefront)."""
    rng = np.random.default_rng(21)
    X = np.concatenate([np.full((20000, 2), 0.3, f32),                      # one run, one cell
                        (0.6 + 0.03 * rng.random((15000, 2))).astype(f32),  # ~4 dense cells
                        rng.random((5000, 2)).astype(f32)])
    X = X[rng.permutation(len(X))]
    _say("iteration many-chunk cells", X.shape)
    a = S.iteration(X, 0.05)
    b = O.egg_iteration(X, 0.05)
    for k in ("order", "N", "half", "border", "synced"):
        assert np.array_equal(a[k], b[k]), k
    assert (a["flag_local"], a["flag_final"]) == (b["flag_local"], b["flag_final"])
    # Positions against the FP64 oracle pass (equal to an FP64 brute-force
    # update here to 1e-16).  The FP32 oracle adds the 20,000 equal terms of
    # the bitwise-equal run one by one, as the reference does, and is ~1e-5
    # off next to that run; the GPU weights the run's term by its length.  Bar:
    # per point, no further from FP64 than the FP32 restatement, plus 1e-6.
    x64 = O.egg_solve(X, 0.05, max_iter=1, fp64=True)["X"].astype(np.float64)
    e_gpu = np.abs(a["X_next"].astype(np.float64) - x64).max(axis=1)
    e_o32 = np.abs(b["X_next"].astype(np.float64) - x64).max(axis=1)
    _say("  max |x - x_fp64|: gpu", e_gpu.max(), "fp32 oracle", e_o32.max())
    assert np.all(e_gpu <= e_o32 + 1e-6), float((e_gpu - e_o32).max())


def _cases_solve():
    from egg_sync import data
    return [("c0", *data.config0()),
            ("c0_eps02", data.config0()[0], 0.02),
            ("c1_20k", data.config1(n=20000)[0], 0.02),
            ("c2_5k", data.config2(n=5000)[0], 0.1),
            ("d1", data.gaussian_blobs(2000, 1, 5, seed=9, center_lo=0.1, center_hi=0.9,
                                       sigma=0.02, noise=0.05), 0.02),
            ("d5", data.gaussian_blobs(3000, 5, 6, seed=11, center_lo=0.2, center_hi=0.8,
                                       sigma=0.05, noise=0.05), 0.15),
            ("d12", data.gaussian_blobs(3000, 12, 5, seed=12, center_lo=0.2, center_hi=0.8,
                                        sigma=0.03, noise=0.05), 0.3),
            ("d20", data.gaussian_blobs(3000, 20, 5, seed=20, center_lo=0.2, center_hi=0.8,
                                        sigma=0.03, noise=0.05), 0.3)]


@pytest.mark.parametrize("case", _cases_solve(), ids=lambda c: c[0])
def test_solve_matches_oracle(S, O, case):
    name, X, eps = case
    _say("solve", name, X.shape, eps)
    a = S.solve(X, eps)
    b = O.egg_solve(X, eps)
    assert a["stats"]["iterations"] == b["iterations"]
    assert a["stats"]["status"] == 0
    assert O.same_partition(a["labels"], b["labels"])
    assert np.abs(a["X"] - b["X"]).max() <= 1e-5
    assert a["stats"]["n_clusters"] == len(set(b["labels"][b["labels"] >= 0]))


def test_sync_mode_matches_reference_golden(S, O):
    g = golden("c0_sync")
    _say("sync mode c0")
    from egg_sync import _native as nat
    a = S.solve(g["X"], float(g["eps"]), stop_rule=nat.STOP_RC, lam=float(g["lam"]))
    assert a["stats"]["iterations"] == int(g["iterations"])
    assert O.same_partition(a["labels"], g["labels"])
    assert np.array_equal(a["labels"], g["labels"])  # same greedy numbering too
    assert np.abs(a["X"] - g["X_final"]).max() <= 1e-6
    # r_c of the last pass is taken on its pre-update positions (SynC.cpp:285):
    # FP64 sums here vs the oracle's FP64 r_c of the same positions; the
    # reference's own value is an FP32 running sum over n points
    it = int(g["iterations"])
    pre = O.sync_solve(g["X"], float(g["eps"]), g["lam"], max_iter=it - 1)["X"]
    assert abs(a["stats"]["rc"] - O.rc(pre, float(g["eps"]))) <= 1e-7
    assert abs(a["stats"]["rc"] - float(g["r_trace"][-1])) <= 2e-6 * np.sqrt(len(g["X"])) / 10


def test_rc_matches_oracle(S, O):
    from egg_sync import data
    X, eps = data.config0()
    Xk = _advance(O, X, eps, 3)
    for Y in (X, Xk):
        assert abs(S.rc(Y, eps) - O.rc(Y, eps)) <= 1e-7


def test_edge_cases(S, O):
    from egg_sync import _native as nat
    _say("edge cases")
    r = S.solve(np.array([[0.25, 0.75]], f32), 0.05)
    assert list(r["labels"]) == [-1] and r["stats"]["iterations"] == 1
    X = np.tile(np.array([[0.4, 0.6, 0.1]], f32), (9, 1))
    r = S.solve(X, 0.05)
    assert len(set(r["labels"])) == 1 and r["labels"][0] >= 0
    X = np.array([[0.0, 0.0], [1.0, 1.0], [0.0, 1.0], [1.0, 0.0], [1.0, 1.0]], f32)
    a, b = S.solve(X, 0.05), O.egg_solve(X, 0.05)
    assert O.same_partition(a["labels"], b["labels"])
    # eps so large that the grid is a single cell (w clamps to 1)
    rng = np.random.default_rng(1)
    X = rng.random((200, 2)).astype(f32)
    a, b = S.solve(X, 3.0), O.egg_solve(X, 3.0)
    assert a["stats"]["iterations"] == b["iterations"]
    assert O.same_partition(a["labels"], b["labels"])
    # iteration cap comes back as a status, not an exit
    from egg_sync import data
    Xc, eps = data.config0()
    r = S.solve(Xc, eps, max_iter=1)
    assert r["code"] == nat.CAPPED and r["stats"]["status"] == 1 and r["stats"]["iterations"] == 1
    # invalid input
    for bad in (np.array([[0.5, 1.5]], f32), np.array([[np.nan, 0.1]], f32),
                np.array([[-0.1, 0.1]], f32)):
        with pytest.raises(nat.EggSyncError) as e:
            S.solve(bad, 0.05)
        assert e.value.code == nat.E_RANGE
    with pytest.raises(nat.EggSyncError):
        S.solve(np.zeros((4, 21), f32), 0.05)


def test_deterministic(S):
    from egg_sync import data
    X, eps = data.config1(n=50000)
    a, b = S.solve(X, eps), S.solve(X, eps)
    assert np.array_equal(a["X"], b["X"]) and np.array_equal(a["labels"], b["labels"])
    assert a["stats"]["iterations"] == b["stats"]["iterations"]


def test_torch_extension_api(S, O):
    from egg_sync import algorithms, data
    X, eps = data.config0()
    _say("torch extension")
    for D in (torch.from_numpy(X), torch.from_numpy(X).cuda()):
        labels, itr_times, stage_times, space = algorithms.EGG_SynC(D, eps)
        assert labels.dtype == torch.float32 and labels.shape == (len(X),)
        assert stage_times.shape == (6,) and space.numel() == 1 and int(space[0]) > 0
        ref = O.egg_solve(X, eps)
        assert itr_times.shape == (ref["iterations"],)
        assert O.same_partition(labels.cpu().numpy().astype(np.int64), ref["labels"])
    g = golden("c0_sync")
    labels, itr_times, _, _ = algorithms.SynC(torch.from_numpy(g["X"]), float(g["eps"]))
    assert labels.dtype == torch.int32
    assert np.array_equal(labels.numpy(), g["labels"])
    assert itr_times.shape == (int(g["iterations"]),)


def test_fsync_matches_reference_fsync(S):
    """FSynC(D, eps) (python/algorithms.py:98-99) against the labels and
    iteration count of the reference's own FDynamicalClustering on config 0
    (tests/golden/c0_fsync.npz, written by the compiled reference)."""
    from egg_sync import algorithms
    g = golden("c0_fsync")
    _say("FSynC config 0")
    labels, itr_times, stage_times, space = algorithms.FSynC(torch.from_numpy(g["X"]),
                                                             float(g["eps"]), B=int(g["B"]))
    assert labels.dtype == torch.int32
    assert np.array_equal(labels.numpy(), g["labels"])
    assert itr_times.shape == (int(g["iterations"]),)
    labels2, _, _, _ = algorithms.FDynamicalClustering(torch.from_numpy(g["X"]).cuda(),
                                                       float(g["eps"]), B=50)
    assert np.array_equal(labels2.cpu().numpy(), g["labels"])
    with pytest.raises(Exception):
        algorithms.FSynC(torch.from_numpy(g["X"]), float(g["eps"]), B=1)


def test_full_size_config1_properties(S, O):
    """n = 1,000,000, d = 3, eps = 0.02: the oracle's full-size solve (fixture
    written by tests/golden/make_oracle_fixtures.py), determinism, and a
    bit-exact pass of the converged state against the oracle at full size."""
    from egg_sync import data
    X, eps = data.config1()
    _say("config1 full size solve")
    a = S.solve(X, eps)
    assert a["stats"]["status"] == 0
    _say("config1 iterations", a["stats"]["iterations"], "clusters", a["stats"]["n_clusters"])
    ref = golden("c1_full_oracle")
    assert a["stats"]["iterations"] == int(ref["iterations"])
    assert O.same_partition(a["labels"], ref["labels"])
    assert np.abs(a["X"][ref["sample_idx"]] - ref["sample_X"]).max() <= 1e-5
    assert np.abs(a["X"].astype(np.float64).sum(0) - ref["X_sum"]).max() <= 1e-5 * len(X)
    b = S.solve(X, eps)
    assert np.array_equal(a["X"], b["X"]) and np.array_equal(a["labels"], b["labels"])
    # labels are the cell rule applied to the final positions
    assert O.same_partition(a["labels"], O.egg_labels(a["X"], eps))
    # one more pass from the final state: GPU and oracle agree bit for bit
    _say("config1 pass on the final state")
    ga = S.iteration(a["X"], eps)
    gb = O.egg_iteration(a["X"], eps)
    for k in ("order", "N", "half", "border", "synced"):
        assert np.array_equal(ga[k], gb[k]), k
    assert (ga["flag_local"], ga["flag_final"]) == (gb["flag_local"], gb["flag_final"])
    assert np.abs(ga["X_next"] - gb["X_next"]).max() <= 1e-6


def test_full_size_config1_against_fp64_oracle(S, O):
    """n = 1,000,000, d = 3, eps = 0.02 against the oracle run entirely in FP64
    (tests/golden/c1_fp64_oracle.npz): iterations, partition, positions, and
    the FP64 order parameter r_c of the final state."""
    from egg_sync import data
    X, eps = data.config1()
    g = golden("c1_fp64_oracle")
    _say("config1 solve with r_c")
    a = S.solve(X, eps, compute_rc=True)
    assert a["stats"]["iterations"] == int(g["iterations"])
    assert O.same_partition(a["labels"], g["labels"])
    err = float(np.abs(a["X"][g["sample_idx"]].astype(np.float64) - g["sample_X"]).max())
    drc = abs(a["stats"]["rc"] - float(g["rc64"]))
    _say(f"config1: max |x - x_oracle64| {err:.3g}, |rc - rc_oracle64| {drc:.3g}")
    assert err <= 1e-5 and drc <= 1e-6


def test_full_size_config4(S, O):
    """n = 4,000,000, d = 3, eps = 0.01 against the oracle's full solve
    (tests/golden/c4_full_oracle.npz, 24 passes) and the oracle run from the
    GPU's own states.

    The GPU solve is deterministic and converges after 20 passes to 10,344
    clusters, the oracle's count; its partition differs from the oracle's in
    a few dozen of the 4M points.  The difference comes from the dense passes'
    FP32 sums, not from a pass that decides differently: from any state both
    take the same integer decisions (sorted order, neighbour counts, eps/2
    counts, border and sync flags, both flags of the pass), and the oracle
    continued from the GPU's state after 14 passes converges after the same
    total number of passes, to the same partition.  The FP32 oracle adds the
    points of a run of equal positions one by one (its collapsed clusters
    hold up to ~10^5 points), the GPU weights the run's term by its length;
    against an FP64 pass the GPU is the closer of the two."""
    from egg_sync import data
    from egg_sync.metrics import disagreement
    X, eps = data.config4()
    g = golden("c4_full_oracle")
    n = len(X)
    _say("config4 full solve")
    a = S.solve(X, eps)
    it = a["stats"]["iterations"]
    assert a["stats"]["status"] == 0
    b = S.solve(X, eps)
    assert np.array_equal(a["X"], b["X"]) and np.array_equal(a["labels"], b["labels"])
    dis = disagreement(a["labels"], g["labels"])
    _say("config4: gpu", it, "passes,", a["stats"]["n_clusters"], "clusters; oracle",
         int(g["iterations"]), "passes; disagreement", dis)
    assert a["stats"]["n_clusters"] == int(g["labels"].max()) + 1
    assert dis["noise_differs"] + dis["unmatched"] <= 1e-4 * n
    assert O.same_partition(a["labels"], O.egg_labels(a["X"], eps))

    def pass_equal(Xk, tag):
        ga = S.iteration(Xk, eps)
        gb = O.egg_iteration(Xk, eps)
        for k in ("order", "N", "half", "border", "synced"):
            assert np.array_equal(ga[k], gb[k]), (tag, k)
        assert (ga["flag_local"], ga["flag_final"]) == (gb["flag_local"], gb["flag_final"]), tag
        return ga, gb

    # after 13 passes (clusters collapsing into long runs): integer outputs
    # bit-exact; positions per point no further from an FP64 pass than the
    # FP32 oracle's, plus 1e-6
    X13 = S.solve(X, eps, max_iter=13)["X"]
    _say("config4 pass from the state after 13 passes")
    ga, gb = pass_equal(X13, "13")
    x64 = O.egg_solve(X13, eps, max_iter=1, fp64=True)["X"].astype(np.float64)
    e_gpu = np.abs(ga["X_next"].astype(np.float64) - x64).max(axis=1)
    e_o32 = np.abs(gb["X_next"].astype(np.float64) - x64).max(axis=1)
    _say("  max |x - x_fp64|: gpu", e_gpu.max(), "fp32 oracle", e_o32.max())
    assert np.all(e_gpu <= e_o32 + 1e-6), float((e_gpu - e_o32).max())
    # the oracle continued from the state after 14 passes
    cur = S.solve(X, eps, max_iter=14)["X"]
    pass_equal(cur, "14")
    order, more = None, 0
    while more < 3 * it:
        r = O.egg_iteration(cur, eps, order)
        cur, order, more = r["X_next"], r["order"], more + 1
        if r["flag_final"]:
            break
    _say("  oracle from the gpu state after 14 passes: converges after", 14 + more)
    assert 14 + more == it
    assert O.same_partition(O.egg_labels(cur, eps), a["labels"])
    # the converged state
    pass_equal(a["X"], "final")


def test_config1_sample_solve_matches_oracle(S, O):
    from egg_sync import data
    X, eps = data.config1(n=100_000)
    _say("config1 n=100k solve")
    a = S.solve(X, eps)
    b = O.egg_solve(X, eps)
    assert a["stats"]["iterations"] == b["iterations"]
    assert O.same_partition(a["labels"], b["labels"])
    assert np.abs(a["X"] - b["X"]).max() <= 1e-5


def test_full_size_config2_against_fp32_and_fp64_oracle(S, O):
    """n=250,000, d=8, eps=0.1 (od=3: 175,616 outer cells) against the
    committed oracle runs: FP32 exact-mimic (labels, iterations, positions)
    and FP64 (positions, and r_c computed entirely in FP64)."""
    from egg_sync import data
    g = golden("c2_full_oracle")
    X, eps = data.config2()
    _say("config2 full", X.shape, eps)
    a = S.solve(X, eps, compute_rc=True)
    _say("config2 iterations", a["stats"]["iterations"], "rc", a["stats"]["rc"])
    assert a["stats"]["iterations"] == int(g["iterations"])
    assert O.same_partition(a["labels"], g["labels"])
    assert O.same_partition(a["labels"], g["labels64"])
    idx = g["sample_idx"]
    err32 = float(np.abs(a["X"][idx] - g["sample_X"]).max())
    err64 = float(np.abs(a["X"][idx].astype(np.float64) - g["sample_X64"]).max())
    drc = abs(a["stats"]["rc"] - float(g["rc64"]))
    _say(f"config2: max |x - x_oracle32| {err32:.3g}, max |x - x_oracle64| {err64:.3g}, "
         f"|rc - rc_oracle64| {drc:.3g}")
    assert err32 <= 1e-5 and err64 <= 1e-5
    assert np.abs(a["X"].astype(np.float64).sum(0) - g["X_sum"]).max() <= 1e-5 * len(X)
    assert drc <= 1e-6


@pytest.mark.parametrize("name", ["banknote", "Yeast", "Wilt", "CCPP", "pendigits"])
def test_real_datasets_match_oracle(S, O, name):
    """Real data sets of the reference's run_real (run_experiment.py:53-63) and
    pendigits (17 columns), min-max normalised, eps = 0.05 (run_experiment.py:12),
    against the committed oracle solves (tests/golden/make_real_fixtures.py)."""
    g = golden("real_oracle")
    X, eps = g[f"{name}_X"], float(g["eps"])
    _say("real data set", name, X.shape, eps)
    a = S.solve(X, eps)
    assert a["stats"]["iterations"] == int(g[f"{name}_iterations"])
    assert O.same_partition(a["labels"], g[f"{name}_labels"])
    err = float(np.abs(a["X"] - g[f"{name}_final"]).max())
    _say(f"  max |x - x_oracle| {err:.3g}")
    if err > 1e-5:
        # a pair at distance eps +- 1 ulp took the other branch somewhere on
        # the way (same partition, another valid trajectory): every GPU pass
        # must then equal the oracle's pass from the same GPU state
        it = int(g[f"{name}_iterations"])
        prev = X
        for t in range(1, it + 1):
            ga = S.iteration(prev, eps)
            gb = O.egg_iteration(prev, eps)
            for k in ("order", "N", "half", "border", "synced"):
                assert np.array_equal(ga[k], gb[k]), (t, k)
            assert np.abs(ga["X_next"] - gb["X_next"]).max() <= 1e-6, t
            prev = S.solve(X, eps, max_iter=t)["X"]


def test_real_harness_reads_data_files(tmp_path):
    """experiments.run_real (run_experiment.py:547-551) on a data directory holding
    one file in the reference's format."""
    from egg_sync import experiments as E
    g = golden("real_oracle")
    X = g["banknote_X"].astype(np.float64)
    lab = np.zeros((len(X), 1))
    np.savetxt(tmp_path / "data_banknote_authentication.txt", np.hstack([X, lab]),
               delimiter=",", fmt="%.9g")
    recs = E.run_real(str(tmp_path), rounds=2)
    assert [r["dataset"] for r in recs] == ["banknote", "banknote"]
    assert all(r["iterations"] == int(g["banknote_iterations"]) for r in recs)
    assert list(recs[0]["stage_times"]) == E.STAGES
