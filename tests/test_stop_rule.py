"""Pins EGG-SynC's local-sync stop rule in the CPU oracle.

The rule (src/algorithms/EGG_SynC.cu:583-597, 612-636, 640-709) decides every
iteration count of the solve.  Three independent checks of the oracle's
restatement (oracle/egg_core.inc, update_all):

1. `egg_flags_literal`: a second restatement written straight from the
   reference's source, with none of the oracle's devices -- FP32 box
   distances and FP32 `sqrtf` compares instead of squared-distance
   thresholds, the on-border test in double, the eps/2 compare in double,
   one brute-force pass over all points instead of runs or cell order.  The
   oracle's per-point N, half, border and synced must equal it on the golden
   step inputs, on late states of solves, and on inputs with points clamped
   into cell w-1 and points on and near cell faces.
2. Hand-built configurations whose N / half / border / synced are worked out
   from the reference's source in the comments: a neighbour at exactly eps/2,
   the first float beyond it, the 1e-5 border band, an eps/2-neighbour in a
   cell the box test classes partial, border points alone in their cell.
3. A closed form: two points in one dimension at separation delta < eps keep
   their midpoint and contract to delta - sin(delta) per pass, so the pass at
   which both become locally synced -- and hence the iteration count of the
   whole solve -- follows from the exact trajectory.
"""
import math

import numpy as np
import pytest

from conftest import golden

f32 = np.float32


# ------------------------------------------------- literal restatement ----

def egg_geometry_literal(n, d, eps):
    """EGG_SynC.cu:1249-1260 and :563, as written (FP64 -> FP32 where the source converts)."""
    e = f32(eps)
    cs = f32(math.sqrt(math.pow(float(e) / 2.0, 2.0) / d))     # :1249
    w = int(f32(1.0) / cs)                                       # :1250, float division
    od = min(d, int(math.log(n) / math.log(w)))                  # :1256
    R = int(math.ceil(float(e / cs)))                            # :563, ceil(eps / cell size)
    return e, cs, w, od, R


def egg_flags_literal(X, eps):
    """N, half, border, synced of every point (point-id order) for one pass of
    kernel_itr_grid_STAD5_5_1 (EGG_SynC.cu:551-709)."""
    X = np.ascontiguousarray(X, f32)
    n, d = X.shape
    e, cs, w, od, R = egg_geometry_literal(n, d, eps)
    assert od >= 1, "inputs of these tests keep the reference's outer grid defined"
    # cell dim ids (:103-107): FP32 division, truncation toward zero, w -> w-1
    g = np.trunc(X / cs).astype(np.int64)
    g[g == w] = w - 1
    cells, cell_of = np.unique(g, axis=0, return_inverse=True)
    cell_of = cell_of.reshape(-1)
    count = np.bincount(cell_of, minlength=len(cells))
    # on-border (:586-595): x -+ 0.00001 in double, x / cs in float, no clamp
    xd = X.astype(np.float64)
    idm = np.trunc((xd - 0.00001) / float(cs))
    idp = np.trunc((xd + 0.00001) / float(cs))
    idc = np.trunc(X / cs).astype(np.float64)
    border = np.any((idc != idm) | (idc != idp), axis=1)
    lo_all = cells.astype(f32) * cs                              # dim_id * inner_cell_size
    hi_all = (cells + 1).astype(f32) * cs
    N = np.zeros(n, np.int64)
    half = np.zeros(n, np.int64)
    synced = np.zeros(n, bool)
    half_eps = float(e) / 2.0                                    # eps / 2. (double)
    for p in range(n):
        x = X[p]
        # candidates: inner cells of the outer cells within R (:598-606)
        near = np.all(np.abs(cells[:, :od] - g[p, :od]) <= R, axis=1)
        small = np.zeros(len(cells), f32)
        large = np.zeros(len(cells), f32)
        for l in range(d):                                       # :612-634
            a = lo_all[:, l] - x[l]
            b = hi_all[:, l] - x[l]
            a = a * a
            b = b * b
            small = small + np.where(a < b, a, b)
            large = large + np.where(a < b, b, a)
        included = near & (np.sqrt(small) <= e)
        full = near & (np.sqrt(large) <= e)
        s = np.zeros(n, f32)
        for l in range(d):                                       # gpu_distance, :539-549
            diff = x[l] - X[:, l]
            s = s + diff * diff
        dist = np.sqrt(s)
        in_full = full[cell_of]
        in_part = included[cell_of] & ~full[cell_of]
        N[p] = int(in_full.sum()) + int((in_part & (dist <= e)).sum())   # :657, :690
        if border[p]:                                            # :666-681
            own_cell = in_full & (cell_of == cell_of[p])
            near_half = in_full & (cell_of != cell_of[p]) & (dist.astype(np.float64) <= half_eps)
            half[p] = int(own_cell.sum()) + int(near_half.sum())
        own = int(count[cell_of[p]])
        synced[p] = not ((not border[p] and N[p] != own) or (border[p] and N[p] != half[p]))
    return dict(N=N, half=half, border=border.astype(np.int64), synced=synced.astype(np.int64))


def oracle_flags(O, X, eps):
    it = O.egg_iteration(X, eps)
    out = {}
    for k in ("N", "half", "border", "synced"):
        v = np.empty(len(X), np.int64)
        v[it["order"]] = it[k]
        out[k] = v
    out["flag_local"] = it["flag_local"]
    return out


def assert_flags_equal(O, X, eps):
    lit = egg_flags_literal(X, eps)
    ora = oracle_flags(O, X, eps)
    for k in ("N", "border", "synced"):
        bad = np.nonzero(lit[k] != ora[k])[0]
        assert bad.size == 0, (k, bad[:10], lit[k][bad[:10]], ora[k][bad[:10]])
    # half is only defined (and only used) for border points
    b = lit["border"] == 1
    bad = np.nonzero(lit["half"][b] != ora["half"][b])[0]
    assert bad.size == 0, ("half", bad[:10])
    assert ora["flag_local"] == int(lit["synced"].all())
    return lit


# ------------------------------------------- 1. against the oracle --------

@pytest.mark.parametrize("name", ["step_c0", "step_c0_mid", "step_d3", "step_d5", "step_d8"])
def test_flags_golden_inputs(O, name):
    g = golden(name)
    lit = assert_flags_equal(O, g["X"], float(g["eps"]))
    # the inputs exercise both branches of the rule
    assert lit["border"].sum() > 0 and (lit["border"] == 0).sum() > 0


@pytest.mark.parametrize("name,k", [("step_c0", 1), ("step_c0", 2), ("step_c0", 3),
                                    ("step_d3", 2), ("step_d3", 4), ("step_d5", 2)])
def test_flags_late_states(O, name, k):
    """States after k passes: collapsed groups, synced and unsynced points side by side."""
    g = golden(name)
    eps = float(g["eps"])
    Xk = O.egg_solve(g["X"], eps, max_iter=k)["X"]
    lit = assert_flags_equal(O, Xk, eps)
    assert lit["synced"].sum() > 0


def _faces_and_clamp(d, eps, seed):
    """Blobs plus points clamped into cell w-1, on cell faces, and inside and
    outside the 1e-5 border band."""
    from egg_sync import data
    rng = np.random.default_rng(seed)
    X = data.gaussian_blobs(1500, d, 6, seed=seed, center_lo=0.1, center_hi=0.9, sigma=0.02)
    X = X.astype(f32)
    _, cs, w, _, _ = egg_geometry_literal(len(X), d, eps)
    assert f32(w) * cs < f32(1.0)  # a clamp gap (w cs, 1] exists
    idx = rng.permutation(len(X))
    clamp, face, band = idx[:60], idx[60:200], idx[200:300]
    dims = rng.integers(0, d, size=len(X))
    X[clamp, dims[clamp]] = f32(1.0)
    X[clamp[:30], (dims[clamp[:30]] + 1) % d] = f32(1.0)
    k = np.trunc(X[face, dims[face]] / cs).astype(np.int64)
    k = np.clip(k, 1, w - 1)
    faces = (k.astype(f32) * cs).astype(f32)
    ulps = rng.integers(-3, 4, size=len(face)).astype(np.int32)
    X[face, dims[face]] = (faces.view(np.int32) + ulps).view(f32)
    kb = np.clip(np.trunc(X[band, dims[band]] / cs).astype(np.int64), 1, w - 1)
    off = rng.choice([0.99e-5, 1.01e-5, -0.99e-5, -1.01e-5], size=len(band))
    X[band, dims[band]] = (kb.astype(np.float64) * float(cs) + off).astype(f32)
    X = np.clip(X, 0, 1).astype(f32)
    X[clamp[:10]] = X[clamp[10]]          # a run of equal clamped points
    return X


@pytest.mark.parametrize("d,eps,seed", [(2, 0.06, 0), (3, 0.07, 1), (2, 0.0707, 2)])
def test_flags_faces_clamp_band(O, d, eps, seed):
    X = _faces_and_clamp(d, eps, seed)
    lit = assert_flags_equal(O, X, eps)
    assert lit["border"].sum() >= 100
    Xk = O.egg_solve(X, eps, max_iter=2)["X"]
    assert_flags_equal(O, Xk, eps)


# ------------------------------------------- 2. hand-worked cases ---------

def _fill_1d(points):
    """1-D test points plus 36 fillers at cell centres far from them, so that
    n >= w and the reference's outer grid has one dimension (:1256)."""
    cs = 0.03125
    fill = [(c + 0.5) * cs for c in list(range(0, 9)) + list(range(24, 33)) if c < 32]
    fill = fill + fill
    X = np.array([[v] for v in list(points) + fill], f32)
    return X


def test_hand_exact_half_eps_on_border(O):
    # eps = 2^-4, d = 1: cs = sqrt((eps/2)^2) = 2^-5 exactly, w = 32, R = 2.
    # p = 0.5 lies on the face between cells 15 and 16 -> cell 16, border
    # ((0.5 - 1e-5)/cs truncates to 15).  q = 0.53125 = 17 cs, cell 17, border.
    # For p, cell 17 = [17cs, 18cs]: nearest face 2^-5, farthest 2^-4 <= eps:
    # full.  |p - q| = 2^-5 = eps/2 exactly, and "dist <= eps / 2." counts it:
    # N = 2 (own + q), half = 1 (own cell) + 1 (q) = 2  ->  synced.
    eps = 0.0625
    X = _fill_1d([0.5, 0.53125])
    lit = assert_flags_equal(O, X, eps)
    assert [lit[k][0] for k in ("N", "half", "border", "synced")] == [2, 2, 1, 1]
    assert [lit[k][1] for k in ("N", "half", "border", "synced")] == [2, 2, 1, 1]


def test_hand_first_float_beyond_half_eps(O):
    # As above with q one ulp above 0.53125: the cell is still full for p
    # (the box test depends on the cell only), N = 2, but |p - q| = 2^-5 + 2^-24
    # > eps/2, so half = 1 (own cell only)  ->  p not synced.
    eps = 0.0625
    q = float(np.nextafter(f32(0.53125), f32(1)))
    X = _fill_1d([0.5, q])
    lit = assert_flags_equal(O, X, eps)
    assert [lit[k][0] for k in ("N", "half", "border", "synced")] == [2, 1, 1, 0]


def test_hand_border_band(O):
    # The band is 1e-5 in double: 0.5 + 0.99e-5 is on the border of cell 16
    # ((x - 1e-5)/cs truncates to 15), 0.5 + 1.01e-5 is not.  Each point is
    # alone in cell 16 together with the other one (both in [0.5, 0.53125)):
    #   a = 0.50000990: border, own cell full and holds a and b -> half = 2,
    #       N = 2 -> synced
    #   b = 0.50001010: not border, N = 2 = own -> synced
    # and a third point c = 0.46876 (cell 15, not border) within eps of both:
    # N(a) = N(b) = 3; a: half = 2 + (c within eps/2 in the full cell 15) = 3 -> synced;
    # b: N = 3 != own 2 -> not synced.
    eps = 0.0625
    a, b, c = 0.5 + 0.99e-5, 0.5 + 1.01e-5, 0.46876
    X = _fill_1d([a, b, c])
    lit = assert_flags_equal(O, X, eps)
    assert lit["border"][0] == 1 and lit["border"][1] == 0 and lit["border"][2] == 0
    assert list(lit["N"][:3]) == [3, 3, 3]
    assert lit["half"][0] == 3
    assert list(lit["synced"][:3]) == [1, 0, 0]


def test_hand_half_neighbour_in_partial_cell(O):
    # eps = 0.05, d = 2: cs = f32(sqrt(0.025^2 / 2)) = 0.017677670, w = 56.
    # p = (8cs, 8cs) sits on a corner of cell (8, 8) (border in both dims).
    # q = (9cs, 9cs) is the corner of cell (9, 9).  The FP32 box test of cell
    # (9, 9) for p gives a farthest distance sqrtf(2 (2cs)^2) = 0.0500000082 >
    # eps: the cell is partial, not full.  |p - q| = 0.0249999929 <= eps/2.
    # N counts q (partial cell, |p - q| <= eps): N = 2.  half counts the own
    # cell and eps/2-neighbours in OTHER FULL cells only: half = 1 -> p is not
    # synced although q is within eps/2.
    eps = 0.05
    cs = f32(math.sqrt(math.pow(float(f32(eps)) / 2.0, 2.0) / 2))
    p = [f32(8) * cs, f32(8) * cs]
    q = [f32(9) * cs, f32(9) * cs]
    fill = [[(c + 0.5) * float(cs), (50 + 0.5) * float(cs)] for c in range(0, 56)]
    X = np.array([p, q] + fill, f32)
    lit = assert_flags_equal(O, X, eps)
    assert [lit[k][0] for k in ("N", "half", "border", "synced")] == [2, 1, 1, 0]


def test_hand_squared_distance_at_threshold(O):
    # eps = 2^-4, d = 2: cs = f32(sqrt((2^-5)^2 / 2)) = 0.022097087, w = 45.
    # p = (11 cs, 0.5) lies on a face in dimension 0 (border).  q =
    # (0.27431795, 0.500007689): FP32 differences dx = -2^-5 - ..., dy =
    # -7.689e-6 give s = fl(fl(dx^2) + fl(dy^2)) = 2^-10 + 2^-33, the largest
    # float whose sqrtf is 2^-5 = eps/2, so "dist <= eps / 2." holds with
    # equality.  q's cell (12, 22) is full for p (farthest corner 0.046).
    # N = 2, half = 1 + 1 = 2 -> synced (a strict compare would leave half = 1).
    eps = 0.0625
    cs = f32(math.sqrt(math.pow(0.03125, 2.0) / 2))
    p = [f32(11) * cs, f32(0.5)]
    q = [f32(0.2743179500102997), f32(0.500007688999176)]
    s = f32(0)
    for l in range(2):
        df = f32(p[l] - q[l])
        s = f32(s + f32(df * df))
    assert s == f32(O.sqrt_threshold(0.03125)) and np.sqrt(s) == f32(0.03125)
    fill = [[(c + 0.5) * float(cs), (40 + 0.5) * float(cs)] for c in range(45)]
    X = np.array([p, q] + fill, f32)
    lit = assert_flags_equal(O, X, eps)
    assert [lit[k][0] for k in ("N", "half", "border", "synced")] == [2, 2, 1, 1]


def test_hand_border_test_in_double(O):
    # x = 0.49998998641967773 lies 1.0014e-5 below the face 0.5 = 16 cs
    # (eps = 2^-4, d = 1, cs = 2^-5).  In double, (x + 1e-5) / cs = 15.9999996
    # truncates to 15 = id: not a border point.  (In float, x + 1e-5 rounds to
    # 0.5 and the point would be on the border.)  With y = 0.48 in the same
    # cell 15 and r = 0.51 in cell 16 (full for x, |x - r| <= eps/2):
    # N = 3 != own = 2 -> not synced.  (As a border point, half = 2 + 1 = 3 =
    # N would make it synced.)
    eps = 0.0625
    X = _fill_1d([0.49998998641967773, 0.48, 0.51])
    lit = assert_flags_equal(O, X, eps)
    assert [lit[k][0] for k in ("N", "border", "synced")] == [3, 0, 0]


def test_hand_border_point_alone(O):
    # A border point alone in its cell with no neighbour: N = 1 (own cell),
    # half = 1 -> synced.  Add a neighbour r within eps but beyond eps/2 in a
    # full cell: N = 2, half = 1 -> not synced.  A non-border point alone in
    # its cell with a neighbour in another cell: N = 2 != own = 1 -> not synced.
    eps = 0.0625
    X = _fill_1d([0.5])
    lit = assert_flags_equal(O, X, eps)
    assert [lit[k][0] for k in ("N", "half", "border", "synced")] == [1, 1, 1, 1]
    X = _fill_1d([0.5, 0.45])        # r = 0.45: cell 14 [0.4375, 0.46875], full for p
    lit = assert_flags_equal(O, X, eps)
    assert [lit[k][0] for k in ("N", "half", "border", "synced")] == [2, 1, 1, 0]
    # 0.45 itself: not border, alone in cell 14, sees p -> N = 2 != 1
    assert [lit[k][1] for k in ("N", "border", "synced")] == [2, 0, 0]


# ------------------------------------------- 3. two-point closed form -----

def _two_point_solve_prediction(eps, m, delta0, cs, max_k=20):
    """Exact trajectory of two points m -+ delta_k/2 (delta_{k+1} = delta_k -
    sin delta_k): the iteration count of the solve and the separation after
    it.  Every border decision is asserted to be at least 1e-6 from the 1e-5
    band's edge; a point within 1e-6 of a face may lie in either cell under
    FP32 rounding, and every such choice must give the same decision."""
    import itertools
    e = float(eps)
    delta = delta0
    for k in range(max_k):
        xs = [m - delta / 2, m + delta / 2]
        assert delta < e - 1e-6
        options = []
        for x in xs:
            c = math.floor(x / cs)
            dlo, dhi = x - c * cs, (c + 1) * cs - x
            dface = min(dlo, dhi)
            assert abs(dface - 1e-5) > 1e-6, "trajectory too close to the border band's edge"
            border = dface < 1e-5
            cand = {c}
            if dface < 1e-6:
                cand.add(c - 1 if dlo < dhi else c + 1)
            options.append([(cc, border) for cc in sorted(cand)])
        outcomes = set()
        for (c0, b0), (c1, b1) in itertools.product(*options):
            same = c0 == c1
            N = 2                               # each sees the other (delta <= eps)
            synced = True
            for b in (b0, b1):
                own = 2 if same else 1
                if b:
                    # the other point lies in a full cell (its far face is within
                    # delta + cs < eps) and, here, within eps/2
                    assert delta < e / 2 - 1e-6
                    synced &= N == own + (0 if same else 1)
                else:
                    synced &= N == own
            outcomes.add(synced)
        assert len(outcomes) == 1, "the decision depends on FP32 rounding"
        if outcomes.pop():
            return k + 1, delta - math.sin(delta)
        delta = delta - math.sin(delta)
    raise AssertionError("no stop within max_k")


def _two_point_input(eps, m, delta0, y_row=1):
    """Two points on a line along dimension 0 at height y_row, plus seven
    fillers at cell centres, alone in their cells and farther than eps_extra
    from every other point (synced from the first pass, no extra-check pairs);
    n = 9 = w keeps the reference's outer grid one-dimensional."""
    e, cs, w, od, R = egg_geometry_literal(9, 2, eps)
    assert w == 9 and od == 1
    c = lambda g: (g + 0.5) * float(cs)
    y = c(y_row)
    pts = [[m - delta0 / 2, y], [m + delta0 / 2, y]]
    pts += [[c(6), y]]
    pts += [[c(gx), c(gy)] for gy in (5, 8) for gx in (0, 3, 6)]
    X = np.array(pts, f32)
    # eps_extra (:1430) of eps = 0.3 is 0.3101
    e2 = 2 * float(e) - float(e) * math.sqrt(15 / 16) + float(e) / 2 - math.sin(float(e) / 2)
    for i in range(len(X)):
        for j in range(len(X)):
            if i != j and {i, j} != {0, 1}:
                assert np.linalg.norm(X[i].astype(float) - X[j].astype(float)) > e2 + 1e-3
    return X, float(cs)


@pytest.mark.parametrize("case", [
    # (m, delta0): midpoint on dimension 0, initial separation
    ("centre_far", None, 0.29),       # start in two cells, after one pass in one
    ("straddle", "face", 0.06),       # after one pass still across a face, 1.8e-5 from it
    ("border_route", "face", 0.04),   # after one pass across a face, both in the 1e-5 band
    ("same_cell", None, 0.02),        # same non-border cell from the start
])
def test_two_points_iteration_count_closed_form(O, case):
    name, where, delta0 = case
    eps = 0.3
    _, cs, _, _, _ = egg_geometry_literal(9, 2, eps)
    cs = float(cs)
    if where == "face":
        m = float(f32(2) * f32(cs))  # the face between cells 1 and 2, as the box test computes it
    else:
        m = 1.5 * cs                # the centre of cell 1
    X, csd = _two_point_input(eps, m, delta0)
    m = (float(X[0, 0]) + float(X[1, 0])) / 2
    d0 = float(X[1, 0]) - float(X[0, 0])
    iters, delta_end = _two_point_solve_prediction(eps, m, d0, csd)
    r = O.egg_solve(X, eps, max_iter=10)  # capped: a wrong rule may never stop
    assert r["status"] == 0 and r["iterations"] == iters, (name, r["iterations"], iters)
    expect = {"centre_far": 2, "straddle": 3, "border_route": 2, "same_cell": 1}[name]
    assert iters == expect
    xf = r["X"].astype(np.float64)
    assert abs((xf[1, 0] - xf[0, 0]) - delta_end) <= 2e-7
    assert abs((xf[1, 0] + xf[0, 0]) / 2 - m) <= 2e-7
    assert np.abs(r["X"][2:] - X[2:]).max() <= 1e-7   # the fillers stay put
    # the oracle's per-pass flags agree with the trajectory: synced exactly on
    # the last pass
    Xk = X
    for k in range(iters):
        it = O.egg_iteration(Xk, eps)
        assert it["flag_local"] == (1 if k == iters - 1 else 0)
        Xk = it["X_next"]
