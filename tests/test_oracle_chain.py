"""Pins the CPU oracle's chain against what the mathematics fixes.

The reference cannot run here (its OpenCL wrapper library is not vendored,
Boost is absent and no OpenCL CPU device exists), and it ships no golden
vectors for this path except the prefix-sum known answer.  So each stage of
the restatement is checked against something independent of its own code:
the scan against its known answer and numpy; the grid against an
independent float64 cell computation and the permutation/stability
invariants; the pair and prediction windows against float64 geometry; the
grid walks against brute force over all hits; the filter against circles
and lines with known radius and impact parameter.
"""
import math

import numpy as np
import pytest

from oracle import bridge
from trax_amd import events as ev

from helpers import UNWRITTEN


@pytest.fixture(scope="module")
def pu50():
    w = ev.config_workload(1, 1)
    b = ev.make_batch(w)
    return w, b, bridge.run_chain(w.seed_config, b)


# -------------------------------------------------------------------- scan
def test_scan_known_answer():
    """trax_test.cpp:437-479: all-ones input of 2^20 elements, last == N - 1."""
    n = 1 << 20
    out = bridge.scan_u32(np.ones(n, np.uint32))
    assert out[-1] == n - 1
    assert np.array_equal(out, np.arange(n, dtype=np.uint32))


def test_scan_wraps_mod_2_32():
    rng = np.random.default_rng(1)
    a = rng.integers(0, 1 << 32, 100_000, dtype=np.uint64).astype(np.uint32)
    want = np.concatenate([[0], np.cumsum(a[:-1], dtype=np.uint64) % (1 << 32)]).astype(np.uint32)
    assert np.array_equal(bridge.scan_u32(a), want)


# -------------------------------------------------------------------- grid
@pytest.mark.parametrize("config_id,n_events", [(0, 1), (1, 2), (3, 1)])
def test_grid_invariants(config_id, n_events):
    w = ev.config_workload(config_id, n_events)
    sc = w.seed_config
    b = ev.make_batch(w)
    o = bridge.run_chain(sc, b, stop_after=0)
    C = sc.cells_per_layer
    grid = o["grid"]
    # cell starts are global sorted indices, non-decreasing, segment bases exact
    assert np.all(np.diff(grid.astype(np.int64)) >= 0)
    L = sc.n_layers
    for e in range(b.n_events):
        for l in range(L):
            seg = e * L + l
            assert grid[seg * C] == b.event_offset[e] + b.layer_offset[seg]
    # the sorted hits are a permutation of the input, stable inside each cell
    oi = o["orig_index"]
    assert np.array_equal(np.sort(oi), np.arange(b.n_hits))
    for k, col in zip(("x", "y", "z", "layer", "det_id", "hit_id", "evt"), b.columns()):
        assert np.array_equal(o[k], col[oi]), k
    ends = np.append(grid[1:], b.n_hits)
    cell_of_sorted = np.repeat(np.arange(grid.size), (ends - grid).astype(np.int64))
    assert cell_of_sorted.size == b.n_hits
    same_cell = cell_of_sorted[1:] == cell_of_sorted[:-1]
    assert np.all(oi[1:][same_cell] > oi[:-1][same_cell])
    # every hit lies in its cell: independent float64 phi, exact float32 z binning
    local = cell_of_sorted % C
    zs, ps = local // (sc.n_phi + 1), local % (sc.n_phi + 1)
    z = o["z"]
    zs_ref = np.floor((z - np.float32(-300.0)) / sc.sector_size_z).astype(np.int64)
    assert np.array_equal(zs, np.clip(zs_ref, 0, sc.n_z))
    phi = np.arctan2(o["y"].astype(np.float64), o["x"].astype(np.float64))
    edge_lo = -math.pi + ps * float(sc.sector_size_phi)
    inside = (phi >= edge_lo - 1e-6) & (phi <= edge_lo + float(sc.sector_size_phi) + 1e-6)
    assert inside.all()


def test_grid_phi_boundaries():
    """phi = +pi (y = +0, x < 0) lands in the padding column nPhi; y = -0 gives -pi, column 0."""
    w = ev.config_workload(1, 1)
    sc = w.seed_config
    cell, phi = bridge.cell_of(sc, -5.0, 0.0, 1.0)
    assert phi == np.float32(math.pi) and cell % (sc.n_phi + 1) == sc.n_phi
    cell, phi = bridge.cell_of(sc, -5.0, -0.0, 1.0)
    assert phi == -np.float32(math.pi) and cell % (sc.n_phi + 1) == 0
    # z outside [-300, 300) is clamped to the padding row
    cell, _ = bridge.cell_of(sc, 5.0, 1.0, 400.0)
    assert cell // (sc.n_phi + 1) == sc.n_z
    cell, _ = bridge.cell_of(sc, 5.0, 1.0, -400.0)
    assert cell // (sc.n_phi + 1) == 0


# ------------------------------------------------------------------ windows
def _acos_d(v):
    return math.acos(v)


@pytest.mark.parametrize("sign_y", [1.0, -1.0])
def test_pair_window_geometry(sign_y):
    """Beam-spot window of PairGeneratorBeamspot.h:110-162 against float64 geometry."""
    w = ev.config_workload(1, 1)
    sc = w.seed_config
    t = sc.triplets[0]
    rng = np.random.default_rng(5)
    r1min, r1max = sc.min_radius[0], sc.max_radius[0]
    R = float(sc.min_radius_curvature)
    sig_phi_pair = np.array([t.pair_spread_phi], np.uint32).view(np.float32)[0]
    for _ in range(300):
        phi0 = rng.uniform(0.2, math.pi - 0.2) * sign_y
        rr = 7.3
        X, Y = np.float32(rr * math.cos(phi0)), np.float32(rr * math.sin(phi0))
        Z = np.float32(rng.uniform(-25, 25))
        win = bridge.pair_window(sc, 0, X, Y, Z)
        r = math.hypot(float(X), float(Y))
        z0 = float(int(t.z0))
        zh = max(z0 + (float(Z) - z0) * r1 / r for r1 in (r1min, r1max)) + t.sigma_z
        zl = min(-z0 + (float(Z) + z0) * r1 / r for r1 in (r1min, r1max)) - t.sigma_z
        assert abs(win["zHigh"] - zh) < 2e-4 and abs(win["zLow"] - zl) < 2e-4
        rs = r if sign_y > 0 else -r  # the reference's signed radius (Q2)
        dphi = max(abs(_acos_d(rs / (2 * R)) - _acos_d(r1 / (2 * R))) for r1 in (r1min, r1max))
        dphi += max(math.atan(t.tip_cut * (rs - r1)) / (rs * r1) for r1 in (r1min, r1max))
        phi = math.atan2(float(Y), float(X))
        lo, hi = phi - dphi - float(sig_phi_pair), phi + dphi + float(sig_phi_pair)
        assert win["wrap"] == int(lo < -math.pi or hi > math.pi)
        if lo < -math.pi:
            lo += 2 * math.pi
        if hi > math.pi:
            hi -= 2 * math.pi
        assert abs(win["phiLow"] - lo) < 2e-5 and abs(win["phiHigh"] - hi) < 2e-5


def test_pair_window_wraps_at_pi():
    w = ev.config_workload(1, 1)
    sc = w.seed_config
    X, Y = np.float32(-7.3), np.float32(0.01)  # phi just below +pi
    win = bridge.pair_window(sc, 0, X, Y, np.float32(0.0))
    assert win["wrap"] == 1
    assert win["phiHigh"] < 0 < win["phiLow"]  # high edge wrapped to -pi + ...
    assert win["phiHighSector"] < win["phiLowSector"]


def test_predict_window_geometry():
    """Straight-line r-z extrapolation of TripletThetaPhiPredictor.h:111-133."""
    w = ev.config_workload(1, 1)
    sc = w.seed_config
    t = sc.triplets[0]
    rng = np.random.default_rng(9)
    minR, maxR = sc.min_radius[2], sc.max_radius[2]
    for _ in range(200):
        phi = rng.uniform(-3, 3)
        cot = rng.uniform(-3, 3)
        zv = rng.uniform(-5, 5)
        p1 = [np.float32(4.4 * math.cos(phi)), np.float32(4.4 * math.sin(phi)), np.float32(zv + 4.4 * cot)]
        p2 = [np.float32(7.3 * math.cos(phi)), np.float32(7.3 * math.sin(phi)), np.float32(zv + 7.3 * cot)]
        for store in (False, True):
            win = bridge.predict_window(sc, 0, store, p1, p2, minR, maxR)
            r1 = math.hypot(float(p1[0]), float(p1[1]))
            r2 = math.hypot(float(p2[0]), float(p2[1]))
            c = (float(p2[2]) - float(p1[2])) / (r2 - r1)
            zs = [float(p1[2]) + (rr - r1) * c for rr in (minR, maxR)]
            tol = 2e-3 * (1 + abs(c))
            assert abs(win["zLow"] - (min(zs) - t.sigma_z)) < tol
            assert abs(win["zHigh"] - (max(zs) + (t.sigma_z if store else t.sigma_phi))) < tol


# --------------------------------------------------------------- brute force
def test_pairs_equal_brute_force(pu50):
    """Grid walk == brute force over every inner hit, for every outer hit."""
    w, b, o = pu50
    sc = w.seed_config
    C = sc.cells_per_layer
    phi = bridge.eval_f32("atan2", o["y"], o["x"])
    pairs = o["pairs"][o["pairs"][:, 0] != UNWRITTEN]
    by_outer = {}
    for j, i in pairs:
        by_outer.setdefault(int(i), set()).add(int(j))
    t = sc.triplets[0]
    g1 = o["grid"][(t.layer1 - 1) * C:(t.layer1) * C]
    g2 = o["grid"][(t.layer2 - 1) * C:(t.layer2) * C]
    inner = np.arange(g1[0], g1[-1])
    n_checked = 0
    for i in range(g2[0], g2[-1]):
        win = bridge.pair_window(sc, 0, o["x"][i], o["y"][i], o["z"][i])
        zi, pi_ = o["z"][inner], phi[inner]
        ok = (win["zLow"] <= zi) & (zi <= win["zHigh"])
        if win["wrap"]:
            ok &= ((win["phiLow"] <= pi_) & (pi_ <= np.float32(math.pi))) | (pi_ <= win["phiHigh"])
        else:
            ok &= (win["phiLow"] <= pi_) & (pi_ <= win["phiHigh"])
        assert set(inner[ok].tolist()) == by_outer.get(i, set()), i
        n_checked += 1
    assert n_checked > 500 and len(pairs) > 1000


def test_candidates_against_brute_force(pu50):
    """Candidates of each pair == the store window's hits, truncated at the
    count window's number (TripletThetaPhiPredictor.h:394)."""
    w, b, o = pu50
    sc = w.seed_config
    C = sc.cells_per_layer
    t = sc.triplets[0]
    phi = bridge.eval_f32("atan2", o["y"], o["x"])
    g3 = o["grid"][(t.layer3 - 1) * C:(t.layer3) * C]
    third = np.arange(g3[0], g3[-1])
    assert o["n_cands_unwritten"] == 0
    cands = o["cands"]
    order = np.argsort(cands[:, 0], kind="stable")
    cs = cands[order]
    bounds = np.searchsorted(cs[:, 0], np.arange(len(o["pairs"]) + 1))
    minR, maxR = sc.min_radius[t.layer3 - 1], sc.max_radius[t.layer3 - 1]
    rng = np.random.default_rng(3)
    pick = rng.choice(len(o["pairs"]), size=min(1500, len(o["pairs"])), replace=False)

    def members(win):
        z3, p3 = o["z"][third], phi[third]
        ok = (win["zLow"] <= z3) & (z3 <= win["zHigh"])
        if win["wrap"]:
            ok &= ((win["phiLow"] <= p3) & (p3 <= np.float32(math.pi))) | (p3 <= win["phiHigh"])
        else:
            ok &= (win["phiLow"] <= p3) & (p3 <= win["phiHigh"])
        return set(third[ok].tolist())

    for q in pick:
        h1, h2 = o["pairs"][q]
        p1 = [o["x"][h1], o["y"][h1], o["z"][h1]]
        p2 = [o["x"][h2], o["y"][h2], o["z"][h2]]
        n_count = len(members(bridge.predict_window(sc, 0, False, p1, p2, minR, maxR)))
        s_store = members(bridge.predict_window(sc, 0, True, p1, p2, minR, maxR))
        mine = cs[bounds[q]:bounds[q + 1], 1].tolist()
        assert set(mine) <= s_store
        assert len(mine) == min(n_count, len(s_store))


# ------------------------------------------------------------------- filter
def _circle_points(R, d0, phi0, radii, cot=0.7, z0=0.0):
    """Points at transverse radii r on a circle of radius R whose closest
    approach to the origin is d0 (d0 = 0: through the origin)."""
    cx, cy = (R + d0) * math.cos(phi0), (R + d0) * math.sin(phi0)
    pts = []
    for r in radii:
        # intersection of |p| = r with |p - c| = R, the branch nearer phi0 + 90 deg
        D = math.hypot(cx, cy)
        a = (r * r - R * R + D * D) / (2 * D)
        h = math.sqrt(max(r * r - a * a, 0.0))
        ux, uy = cx / D, cy / D
        px, py = a * ux - h * uy, a * uy + h * ux
        s = 2 * R * math.asin(min(1.0, math.hypot(px, py) / (2 * R))) if d0 == 0 else r
        pts.append([np.float32(px), np.float32(py), np.float32(z0 + s * cot)])
    return pts


def test_filter_circle_through_origin():
    w = ev.config_workload(1, 1)
    sc = w.seed_config
    for R in (30.0, 87.5, 400.0):
        pts = _circle_points(R, 0.0, 0.4, [4.4, 7.3, 10.2])
        f = bridge.filter_eval(sc, 0, *pts)
        assert abs(f["radius"] - R) / R < 2e-3
        assert f["tip2"] < 1e-3
        assert abs(f["theta_ratio"] - 1) < 2e-3
        assert f["pass"] == (R >= float(sc.min_radius_curvature))


def test_filter_impact_parameter():
    w = ev.config_workload(1, 1)
    sc = w.seed_config
    pts = _circle_points(100.0, 1.0, -1.1, [4.4, 7.3, 10.2])
    f = bridge.filter_eval(sc, 0, *pts)
    assert abs(math.sqrt(f["tip2"]) - 1.0) < 0.02
    assert not f["pass"]  # tip 1 cm > tipCut 0.5 cm


def test_filter_straight_line_is_rejected():
    """Collinear hits: the paraboloid normal has n.z = 0, the circle radius is
    infinite and the closest-approach point NaN, so the TIP cut rejects."""
    w = ev.config_workload(1, 1)
    sc = w.seed_config
    pts = [[np.float32(r * 0.6), np.float32(r * 0.8), np.float32(r)] for r in (4.4, 7.3, 10.2)]
    f = bridge.filter_eval(sc, 0, *pts)
    assert not f["pass"]
    assert abs(f["theta_ratio"] - 1) < 1e-5 and f["dphi"] < 1e-5


# ---------------------------------------------------------- whole-chain checks
def test_pair_sets_do_not_depend_on_work_group_size(pu50):
    w, b, o = pu50
    base = np.sort(o["pairs"].view(np.uint64).ravel())
    for T in (1, 7, 64):
        w.seed_config.threads = T
        try:
            o2 = bridge.run_chain(w.seed_config, b, stop_after=1)
        finally:
            w.seed_config.threads = 256
        assert np.array_equal(np.sort(o2["pairs"].view(np.uint64).ravel()), base)
        assert np.array_equal(o2["pair_offsets"], o["pair_offsets"])


def test_segments_and_layers(pu50):
    w, b, o = pu50
    sc = w.seed_config
    t = sc.triplets[0]
    pairs, trk = o["pairs"], o["tracklets"]
    assert np.all(o["layer"][pairs[:, 0]] == t.layer1) and np.all(o["layer"][pairs[:, 1]] == t.layer2)
    c = o["cands"][o["cands"][:, 0] != UNWRITTEN]
    assert np.all(o["layer"][c[:, 1]] == t.layer3)
    assert o["tracklet_offsets"][-1] == len(trk) and o["cand_offsets"][-1] == len(o["cands"])
    assert np.all(o["layer"][trk[:, 2]] == t.layer3)


def test_single_muons_are_all_found():
    """Ten 1-10 GeV muons on three layers: each track gives exactly one triplet
    whose three hits carry its track id (PhysicsRecord-style validity)."""
    w = ev.config_workload(0, 1)
    b = ev.make_batch(w)
    o = bridge.run_chain(w.seed_config, b)
    trk = o["tracklets"]
    hid = o["hit_id"][trk]
    assert len(trk) == 10
    assert np.all(hid[:, 0] == hid[:, 1]) and np.all(hid[:, 1] == hid[:, 2])
    assert sorted(hid[:, 0].tolist()) == list(range(1, 11))


def test_pu200_results_do_not_depend_on_z_binning():
    """The high-pile-up workload's pairs, candidates and triplets are the same
    at 64 and 256 z bins (256 is the autotuned default): the z binning only
    changes how many hits each window walk visits."""
    def run(n_z):
        w = ev.config_workload(3, 1)
        w.seed_config.n_z = n_z
        b = ev.make_batch(w)
        o = bridge.run_chain(w.seed_config, b)
        oi = o["orig_index"]
        pairs = {tuple(r) for r in oi[o["pairs"]].tolist()}
        trk = {tuple(r) for r in oi[o["tracklets"]].tolist()}
        return len(o["pairs"]), len(o["cands"]), pairs, trk
    a, b = run(64), run(256)
    assert a[0] == b[0] and a[1] == b[1]
    assert a[2] == b[2] and a[3] == b[3] and len(a[3]) > 1000


# ------------------------------------------- batch = events alone, segments
@pytest.mark.parametrize("config_id,n_events", [(2, 6), (3, 2)])
def test_batch_equals_events_alone(config_id, n_events):
    """The oracle on an event group gives, per event, what it gives on each
    event alone (every (event, layer) segment is processed on its own), so
    a batch can be checked against the oracle event by event."""
    from helpers import compare_batch_per_event, oracle_per_event
    w = ev.config_workload(config_id, n_events)
    b = ev.make_batch(w)
    whole = bridge.run_chain(w.seed_config, b)
    compare_batch_per_event(whole, b, len(w.seed_config.triplets), oracle_per_event(w.seed_config, b))


@pytest.mark.parametrize("config_id,n_events", [(3, 2), (2, 4)])
def test_segment_rows_belong_to_their_segment(config_id, n_events):
    """Segment k = (event e, layer triplet p) of the pairs, the candidates and
    the triplets holds exactly the rows whose first hit lies in event e on
    layer p + 1 (and whose other hits lie in e on the triplet's layers):
    every row of [off[k], off[k+1]) satisfies that, and off[k+1] - off[k]
    equals the number of such rows in the whole array."""
    w = ev.config_workload(config_id, n_events)
    sc = w.seed_config
    b = ev.make_batch(w)
    o = bridge.run_chain(sc, b)
    P, E = len(sc.triplets), b.n_events
    lay, evt = o["layer"].astype(np.int64), o["evt"].astype(np.int64)
    l2 = np.array([t.layer2 for t in sc.triplets])
    l3 = np.array([t.layer3 for t in sc.triplets])
    cands = o["cands"]
    written = cands[:, 0] != UNWRITTEN
    cand_rows = np.zeros((len(cands), 3), np.int64)
    cand_rows[written, :2] = o["pairs"][cands[written, 0]]
    cand_rows[written, 2] = cands[written, 1]
    stages = {"pairs": (o["pairs"].astype(np.int64), o["pair_offsets"], np.ones(len(o["pairs"]), bool)),
              "cands": (cand_rows, o["cand_offsets"], written),
              "tracklets": (o["tracklets"].astype(np.int64), o["tracklet_offsets"],
                            np.ones(len(o["tracklets"]), bool))}
    for name, (rows, off, ok) in stages.items():
        assert off[0] == 0 and off[-1] == len(rows) and np.all(np.diff(off.astype(np.int64)) >= 0), name
        seg = np.searchsorted(off, np.arange(len(rows)), side="right") - 1
        e_of, p_of = seg // P, seg % P
        r, e_r, p_r = rows[ok], e_of[ok], p_of[ok]
        assert np.all(evt[r[:, 0]] == e_r) and np.all(lay[r[:, 0]] == p_r + 1), name
        assert np.all(evt[r[:, 1]] == e_r) and np.all(lay[r[:, 1]] == l2[p_r]), name
        if r.shape[1] > 2:
            assert np.all(evt[r[:, 2]] == e_r) and np.all(lay[r[:, 2]] == l3[p_r]), name
        # every segment's size equals the number of rows that belong to it
        # (unwritten candidate slots stay with the pair that owns them)
        owner = rows[:, 0] if name != "cands" else np.where(written, cand_rows[:, 0], -1)
        for k in range(E * P):
            e, p = divmod(k, P)
            if name == "cands":
                # a segment's slots are the budgets of its own pairs
                pk = slice(int(o["pair_offsets"][k]), int(o["pair_offsets"][k + 1]))
                own = np.isin(cands[:, 0], np.arange(pk.start, pk.stop)) | \
                    ((~written) & (seg == k))
                assert own.sum() == off[k + 1] - off[k], (name, k)
            else:
                belongs = (evt[owner] == e) & (lay[owner] == p + 1)
                assert belongs.sum() == off[k + 1] - off[k], (name, k)
        assert len(r) > 0, name
