"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

- CPU SynC restatement: bitwise against golden vectors written by the
  reference's own src/algorithms/SynC.cpp (tests/golden/make_golden.py).
- EGG one-pass update: against the reference's brute-force pass (same map
  x -> x + mean_{N_eps} sin(x_j - x)) on every point whose eps-ball is not
  affected by the two EGG box-test quirks reproduced from EGG_SynC.cu:612-634;
  for the affected points the neighbour-count difference must be exactly the
  signed number of quirk pairs.
- Grid build: invariants and brute-force grouping.
- Closed forms: single point, identical points, two-point contraction
  d' = d - sin d with fixed midpoint, symmetric triplet, independent clusters.
- Extra check: hand-built configurations in the (eps, eps_extra] band.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden

f32 = np.float32


def _dist2_f32(X, x):
    """Row-wise FP32 squared distance, products and sums rounded separately."""
    s = np.zeros(X.shape[0], f32)
    for l in range(X.shape[1]):
        df = (x[l] - X[:, l]).astype(f32)
        s = (s + (df * df).astype(f32)).astype(f32)
    return s


def _coords(X, cs, w):
    q = (X.astype(f32) / f32(cs)).astype(f32)
    return np.clip(np.trunc(q), 0, w - 1).astype(np.int64)


# ------------------------------------------------------------ CPU SynC ----

def test_sync_oracle_bitwise_against_reference(O):
    g = golden("c0_sync")
    r = O.sync_solve(g["X"], float(g["eps"]), g["lam"])
    assert r["iterations"] == int(g["iterations"])
    assert np.array_equal(r["labels"], g["labels"])
    assert np.array_equal(r["X"], g["X_final"])
    assert np.array_equal(r["r_trace"], g["r_trace"])


def test_reference_fsync_equals_reference_sync_config0():
    # the reference's FSynC (R-tree neighbour search, SynC.cpp:426-487) and SynC
    # (scan, SynC.cpp:268-326) give the same labels and iteration count on
    # config 0: the neighbour sets are equal, only their order differs
    a, b = golden("c0_fsync"), golden("c0_sync")
    assert np.array_equal(a["X"], b["X"]) and float(a["eps"]) == float(b["eps"])
    assert int(a["iterations"]) == int(b["iterations"])
    assert np.array_equal(a["labels"], b["labels"])


@pytest.mark.parametrize("name", ["step_c0", "step_c0_mid", "step_d3", "step_d5", "step_d8"])
def test_sync_single_pass_bitwise(O, name):
    g = golden(name)
    X = g["X"]
    Xn, s = O.sync_iteration_range(X, float(g["eps"]), 0, X.shape[0], nthreads=2)
    assert np.array_equal(Xn, g["X_next"])
    assert abs(s - float(np.sum(g["r_p"].astype(np.float64)))) < 1e-6 * X.shape[0]


def test_sync_labels_reference_semantics(O):
    # synCluster_v2 (SynC.cpp:88-129): greedy in index order with overwrites
    X = np.array([[0.10, 0.5], [0.13, 0.5], [0.16, 0.5], [0.8, 0.8]], f32)
    lab = O.sync_labels(X, 0.035)
    # i=0 labels 1 (cl 0); i=2 is unlabelled: it relabels 1 and takes cl 1
    assert list(lab) == [0, 1, 1, -1]


def test_simple_labels_known_answers(O):
    # GPU_synCluster (simple_GPU_SynC.cu:70-95): smallest eps-neighbour index,
    # not transitive; rename_cluster numbers the representatives in index order
    X = np.array([[0.0], [0.01], [0.5], [0.51], [0.9]], f32)
    lab, k = O.simple_labels(X, 0.05)
    assert list(lab) == [0, 0, 1, 1, -1] and k == 2
    chain = np.array([[0.0], [0.04], [0.08]], f32)
    lab, k = O.simple_labels(chain, 0.05)
    assert list(lab) == [0, 0, 1] and k == 2      # point 2's smallest neighbour is 1
    lab, k = O.simple_labels(np.full((4, 3), 0.3, f32), 0.01)
    assert list(lab) == [0, 0, 0, 0] and k == 1
    lab, k = O.simple_labels(np.array([[0.2, 0.2]], f32), 0.1)
    assert list(lab) == [-1] and k == 0


def _simple_labels_numpy(X, eps):
    """The same rule, vectorised: FP32 squared distances summed dimension by
    dimension, sqrt rounded to FP32."""
    X = X.astype(f32)
    s = np.zeros((len(X), len(X)), f32)
    for l in range(X.shape[1]):
        df = X[:, None, l] - X[None, :, l]
        s = s + df * df
    nb = np.sqrt(s) <= f32(eps)
    cnt = nb.sum(1)
    first = nb.argmax(1)
    reps = np.zeros(len(X), bool)
    reps[first[cnt > 1]] = True
    rank = np.cumsum(reps) - 1
    return np.where(cnt > 1, rank[first], -1).astype(np.int32)


@pytest.mark.parametrize("d,eps", [(1, 0.02), (2, 0.05), (3, 0.1), (8, 0.3), (20, 0.5)])
def test_simple_labels_against_numpy(O, d, eps):
    rng = np.random.default_rng(d)
    X = rng.random((400, d)).astype(f32)
    X[50:60] = X[60]
    lab, k = O.simple_labels(X, eps)
    ref = _simple_labels_numpy(X, eps)
    assert np.array_equal(lab, ref) and k == ref.max() + 1


def test_simple_solve_is_the_sync_pass_with_a_cap(O):
    # kernel_itr (simple_GPU_SynC.cu:29-67) adds the terms of SynC.cpp:16-56 in
    # the same order, so its passes are the compiled reference's (c0_sync)
    g = golden("c0_sync")
    r = O.simple_solve(g["X"], float(g["eps"]), g["lam"])
    it = int(g["iterations"])
    assert r["iterations"] == min(it, 100)
    if it <= 100:
        assert np.array_equal(r["X"], g["X_final"])
    r3 = O.simple_solve(g["X"], float(g["eps"]), lam=np.float32(1.5), max_iter=3)
    assert r3["iterations"] == 3 and r3["status"] == 1


# ------------------------------------------------------------ geometry ----

@pytest.mark.parametrize("n,d,eps,cs,w,od,O_,R", [
    (2000, 2, 0.05, 0.0176777, 56, 1, 56, 3),
    (1_000_000, 3, 0.02, 0.0057735, 173, 2, 29929, 4),
    (250_000, 8, 0.1, 0.0176777, 56, 3, 175616, 6),
    (4_000_000, 3, 0.01, 0.0028868, 346, 2, 119716, 4),
])
def test_geometry_table(O, n, d, eps, cs, w, od, O_, R):
    g = O.geometry(n, d, eps)
    assert abs(g.cs - cs) < 1e-6 and (g.w, g.od, g.O, g.R) == (w, od, O_, R)
    # the cell diagonal is eps/2 (EGG_SynC.cu:1249)
    assert abs(np.sqrt(d) * g.cs - eps / 2) < 1e-7
    assert g.O <= n


def test_eps_extra_value(O):
    g = O.geometry(1_000_000, 3, 0.02)
    assert abs(g.eps_extra - 0.0206352) < 1e-7


@pytest.mark.parametrize("e", [0.05, 0.02, 0.01, 0.1, 0.025, 0.0206352, 1e-3, 0.3333])
def test_sqrt_threshold(O, e):
    e = f32(e)
    t = f32(O.sqrt_threshold(float(e)))
    assert np.sqrt(t) <= e
    assert np.sqrt(np.nextafter(t, f32(np.inf))) > e


# ---------------------------------------------------------- grid build ----

def _bitlen(v):
    return max(1, int(v).bit_length())


@pytest.mark.parametrize("d,n,eps,seed", [(1, 500, 0.05, 0), (2, 2000, 0.05, 1), (3, 3000, 0.05, 2),
                                          (5, 1500, 0.2, 3), (8, 1000, 0.3, 4), (12, 1500, 0.3, 5),
                                          (20, 1500, 0.3, 6), (19, 1000, 0.002, 7)])
def test_grid_invariants(O, d, n, eps, seed):
    rng = np.random.default_rng(seed)
    if d <= 8:
        X = rng.random((n, d)).astype(f32)
    else:
        # one tight cluster around the centre of a cell, so that cells hold
        # many points and their order inside a cell (all d split bits, hash)
        # is exercised
        cs = O.geometry(n, d, eps).cs
        c = (np.floor(0.5 / cs) + 0.5) * cs
        X = np.clip(c + cs / 8 * rng.standard_normal((n, d)), 0, 1).astype(f32)
    X[:5] = 1.0  # exercise the w -> w-1 clamp
    X[5:10] = 0.0
    X[10:20:2] = X[11]  # repeated positions, interleaved with others
    order_in = rng.permutation(n).astype(np.int32)
    G = O.build_grid(X, eps, order_in)
    g = G["geom"]
    order, cell_of, cs_, cg = G["order"], G["cell_of"], G["cell_start"], G["cell_g"]
    assert np.array_equal(np.sort(order), np.arange(n))
    assert cs_[0] == 0 and cs_[-1] == n and np.all(np.diff(cs_) > 0)
    coords = _coords(X, g.cs, g.w)
    assert np.array_equal(coords[order], cg[cell_of])
    assert len({tuple(r) for r in coords}) == G["C"]
    outer = np.zeros(G["C"], np.int64)
    for l in range(g.od):
        outer += cg[:, l] * g.w ** l
    keys = [tuple([o] + list(r[g.od:])) for o, r in zip(outer, cg)]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    # inside a cell, points are ordered by sub-cell (split bits at the cell
    # centre, computed in FP32, dimension 0 first; present while the key
    # stays within 256 bits), then by the low bits of a hash of the
    # position's bit patterns (up to 16, as many as the key's last 64-bit
    # word has room for), and stably within equal keys; bitwise-equal
    # positions therefore sit next to each other
    key_bits = _bitlen(g.O - 1) + (d - g.od) * _bitlen(g.w - 1)
    sub_on = key_bits + d <= 256
    used = key_bits + (d if sub_on else 0)
    hash_bits = min(16, -(-used // 64) * 64 - used)
    if d == 19:
        assert not sub_on and hash_bits == 9
    mid = ((coords.astype(f32) + f32(0.5)) * f32(g.cs)).astype(f32)
    sub = (X >= mid).astype(np.int64)
    sub_key = np.zeros(n, np.int64)
    for l in range(d):
        sub_key = sub_key * 2 + (sub[:, l] if sub_on else 0)
    h = np.full(n, 0x9E3779B9, np.uint64)
    bits = X.view(np.uint32).astype(np.uint64)
    for l in range(d):
        h ^= bits[:, l]
        h = (h * 0x85EBCA6B) & 0xFFFFFFFF
        h ^= h >> 13
    h = (h * 0xC2B2AE35) & 0xFFFFFFFF
    h ^= h >> 16
    full_key = sub_key * (1 << hash_bits) + (h & ((1 << hash_bits) - 1)).astype(np.int64)
    rank = np.empty(n, np.int64)
    rank[order_in] = np.arange(n)
    for c in range(G["C"]):
        pts = order[cs_[c]:cs_[c + 1]]
        fk = full_key[pts]
        assert np.all(np.diff(fk) >= 0)
        for v in np.unique(fk):
            assert np.all(np.diff(rank[pts[fk == v]]) > 0)
    for o in np.unique(outer):
        cells = np.nonzero(outer == o)[0]
        assert G["ob"][o] == cells[0] and G["oe"][o] == cells[-1] + 1
    empty = np.setdiff1d(np.arange(g.O), outer)
    assert np.all(G["ob"][empty] == G["oe"][empty])


# -------------------------------------------- EGG update vs reference ----

def _quirk_pairs(X, eps, g):
    """Signed count per point of pairs where the reference's FP32 box test
    (EGG_SynC.cu:612-634) disagrees with the point's own FP32 distance:
    +1 counted through a full cell although farther than eps, -1 in a cell the
    test rejects although within eps."""
    n, d = X.shape
    cs = f32(g.cs)
    coords = _coords(X, g.cs, g.w)
    thr = f32(g.thr_eps)
    out = np.zeros(n, np.int64)
    lo_all = (coords.astype(f32) * cs).astype(f32)
    hi_all = ((coords + 1).astype(f32) * cs).astype(f32)
    for p in range(n):
        x = X[p]
        sm = np.zeros(n, f32)
        lg = np.zeros(n, f32)
        for l in range(d):
            a = (lo_all[:, l] - x[l]).astype(f32)
            b = (hi_all[:, l] - x[l]).astype(f32)
            a = (a * a).astype(f32)
            b = (b * b).astype(f32)
            sm = (sm + np.where(a < b, a, b)).astype(f32)
            lg = (lg + np.where(a < b, b, a)).astype(f32)
        near = np.all(np.abs(coords[:, :g.od] - coords[p, :g.od]) <= g.R, axis=1)
        within = _dist2_f32(X, x) <= thr
        full = near & (lg <= thr)
        part = near & (sm <= thr) & ~full
        counted = full | (part & within)
        out[p] = int(np.sum(counted & ~within)) - int(np.sum(within & ~counted))
    return out


@pytest.mark.parametrize("name", ["step_c0", "step_c0_mid", "step_d3", "step_d5", "step_d8"])
def test_egg_pass_against_reference_bruteforce(O, name):
    gv = golden(name)
    X, eps = gv["X"], float(gv["eps"])
    n, d = X.shape
    it = O.egg_iteration(X, eps)
    N = np.empty(n, np.int64)
    N[it["order"]] = it["N"]
    quirk = _quirk_pairs(X, eps, O.geometry(n, d, eps))
    assert np.array_equal(N - gv["N"], quirk)
    clean = quirk == 0
    assert clean.mean() > 0.9
    err = np.abs(it["X_next"][clean] - gv["X_next"][clean]).max()
    assert err <= 2e-6, err


def test_egg_pass_against_fp64_bruteforce(O):
    gv = golden("step_d3")
    X, eps = gv["X"], float(gv["eps"])
    n, d = X.shape
    it = O.egg_iteration(X, eps)
    quirk = _quirk_pairs(X, eps, O.geometry(n, d, eps))
    X64 = X.astype(np.float64)
    thr = f32(O.sqrt_threshold(eps))
    for p in np.nonzero(quirk == 0)[0][::25]:
        nb = _dist2_f32(X, X[p]) <= thr
        exact = X64[p] + np.sin(X64[nb] - X64[p]).mean(axis=0)
        assert np.abs(it["X_next"][p] - exact).max() <= 1e-6


# ------------------------------------------------------- closed forms ----

def test_single_point(O):
    X = np.array([[0.37, 0.61]], f32)
    r = O.egg_solve(X, 0.05)
    assert r["iterations"] == 1 and list(r["labels"]) == [-1]
    assert np.array_equal(r["X"], X)


def test_identical_points(O):
    X = np.tile(np.array([[0.31, 0.72, 0.5]], f32), (17, 1))
    r = O.egg_solve(X, 0.05)
    assert r["iterations"] == 1
    assert len(set(r["labels"])) == 1 and r["labels"][0] == 0
    assert np.abs(r["X"] - X).max() <= 2 * np.spacing(f32(0.72))


@pytest.mark.parametrize("x0,delta,eps", [(0.40, 0.012, 0.02), (0.2031, 0.0183, 0.02),
                                          (0.61, 0.03, 0.05)])
def test_two_points_contract_by_sine(O, x0, delta, eps):
    X = np.array([[x0], [x0 + delta]], f32)
    it = O.egg_iteration(X, eps)
    Xn = it["X_next"].astype(np.float64)[:, 0]
    X0 = X.astype(np.float64)[:, 0]
    dd = X0[1] - X0[0]
    assert abs((Xn[1] - Xn[0]) - (dd - np.sin(dd))) <= 4e-7
    assert abs((Xn[1] + Xn[0]) - (X0[1] + X0[0])) <= 4e-7


def test_symmetric_triplet_centre_fixed(O):
    d_ = 2.0 ** -7
    X = np.array([[0.5 - d_], [0.5], [0.5 + d_]], f32)
    it = O.egg_iteration(X, 0.02)
    assert abs(float(it["X_next"][1, 0]) - 0.5) <= 1e-7
    assert abs(float(it["X_next"][0, 0] + it["X_next"][2, 0]) - 1.0) <= 2e-7


def test_independent_clusters(O):
    from egg_sync import data
    A = data.gaussian_blobs(300, 2, 2, seed=5, center_lo=0.2, center_hi=0.8, sigma=0.02)
    A = (A * 0.3 + 0.05).astype(f32)          # inside [0.05, 0.35]^2
    B = (A + f32(0.6)).astype(f32)            # inside [0.65, 0.95]^2, far beyond eps_extra
    U = np.concatenate([A, B])
    eps = 0.03
    ra, rb, ru = O.egg_solve(A, eps), O.egg_solve(B, eps), O.egg_solve(U, eps)
    assert ru["iterations"] == max(ra["iterations"], rb["iterations"])
    assert O.same_partition(ru["labels"][:300], ra["labels"])
    assert O.same_partition(ru["labels"][300:], rb["labels"])
    assert not set(ru["labels"][:300][ru["labels"][:300] >= 0]) & set(ru["labels"][300:])


# --------------------------------------------------------- extra check ----

def test_extra_check_mbr_band(O):
    eps = f32(0.02)
    e = float(eps)
    p = np.array([0.3, 0.3], f32)
    q1 = p + np.array([1.02 * e, 0.0], f32)
    q3 = p + np.array([0.95 * e, 0.4 * e], f32)
    g = O.geometry(3, 2, e)
    assert np.sqrt(_dist2_f32(q1[None], p)[0]) > eps
    assert np.sqrt(_dist2_f32(q1[None], p)[0]) <= f32(g.eps_extra)
    # per-dim nearest coordinates of {q1, q3} to p form (0.95 eps, 0): within eps
    assert O.extra_check(np.stack([p, q1, q3]), e) == 0
    # same p, q1 with q3 pushed out of q1's eps/2-ball: no possible merge
    q3b = p + np.array([0.95 * e, 0.6 * e], f32)
    assert O.extra_check(np.stack([p, q1, q3b]), e) == 1
    # beyond eps_extra: nothing to check
    q4 = p + np.array([1.2 * e, 0.0], f32)
    assert O.extra_check(np.stack([p, q4]), e) == 1


# ------------------------------------------------------------- solves ----

def test_egg_solve_fp32_vs_fp64_config0(O):
    from egg_sync import data
    X, eps = data.config0()
    a = O.egg_solve(X, eps)
    b = O.egg_solve(X.astype(np.float64), eps, fp64=True)
    assert a["iterations"] == b["iterations"]
    assert O.same_partition(a["labels"], b["labels"])
    assert np.abs(a["X"] - b["X"]).max() <= 1e-5


def test_generators_pinned():
    from egg_sync import data
    dg = json.load(open(os.path.join(GOLDEN, "digests.json")))
    assert data.digest(data.config0()[0]) == dg["config0_n2000"]
    assert data.digest(data.config1(n=1000)[0]) == dg["config1_n1000"]
    assert data.digest(data.config2(n=1000)[0]) == dg["config2_n1000"]
    assert [data.digest(x) for x in data.config3_instances(count=3, n_hi=5000)] == dg["config3_first3"]


def test_rc_of_collapsed_clusters(O):
    X = np.repeat(np.array([[0.2, 0.2], [0.7, 0.7]], f32), 50, axis=0)
    assert O.rc(X, 0.05) == 1.0
    X2 = np.array([[0.5, 0.5], [0.5, 0.53]], f32)
    dist = float(np.sqrt(_dist2_f32(X2[1:], X2[0])[0]))
    assert abs(O.rc(X2, 0.05) - (1.0 + float(np.exp(f32(-dist)))) / 2.0) < 1e-7


def test_config1_fp64_fixture_agrees_with_fp32_fixture():
    """The oracle's n=1,000,000 solve run in FP64 and in FP32 (both committed,
    written by tests/golden/make_oracle_fixtures.py): the same iteration count
    and partition, positions within 1e-6 of each other, and an order parameter
    r_c of 1 -- every point's eps-neighbours share its final position, so each
    term exp(-|sin(x_q - x)|) is 1."""
    from egg_sync.metrics import same_partition
    a, b = golden("c1_fp64_oracle"), golden("c1_full_oracle")
    assert int(a["iterations"]) == int(b["iterations"]) == 10
    assert same_partition(a["labels"], b["labels"])
    assert np.array_equal(a["sample_idx"], b["sample_idx"])
    assert np.abs(a["sample_X"] - b["sample_X"].astype(np.float64)).max() <= 1e-6
    assert abs(float(a["rc64"]) - 1.0) <= 1e-12


# ------------------------------------------------- learned-jump MLP oracle ----

def test_mlp_oracle_gradient_finite_differences():
    import mlp_oracle as M
    rng = np.random.default_rng(0)
    n_in, h, n_out, B = 5, 7, 3, 11
    P = rng.normal(0, 0.5, sum(int(np.prod(s)) for _, s in M.sizes(n_in, h, n_out)))
    X = rng.normal(0, 1, (B, n_in))
    T = rng.normal(0, 1, (B, n_out))
    loss, g = M.loss_and_grad(P, X, T, n_in, h, n_out)
    num = np.empty_like(P)
    for k in range(P.size):
        e = np.zeros_like(P)
        e[k] = 1e-6
        num[k] = (M.loss_and_grad(P + e, X, T, n_in, h, n_out)[0] -
                  M.loss_and_grad(P - e, X, T, n_in, h, n_out)[0]) / 2e-6
    assert np.abs(num - g).max() <= 1e-6 * max(1.0, np.abs(g).max())
    O, _ = M.forward(P, X, n_in, h, n_out)
    assert abs(loss - np.sum((O - T) ** 2) / B) < 1e-12
