"""The oracle's PairGeneratorSector restatement (PairGeneratorSector.h:43-234)
pinned without the reference running:

* the window against the cell coordinates computed in float64 (rows
  zs - s .. zs + s, phi cells ps - s .. ps + s, the wrap-around corrections
  and the unsigned underflow of zs - s below row 0);
* the pair set against a brute force over cell membership (an inner hit pairs
  with every outer hit whose cell lies in its row range and phi-cell set),
  which shares nothing with the oracle's index-range walk;
* duplicate visits (a phi window wider than the ring): the bitmap keeps one
  pair, the visit counts (and so the offsets) keep every visit, and the
  overcount equals the closed-form number of repeated visits.
"""
import numpy as np
import pytest

from helpers import rows_sorted, with_segment, written_pairs


def _oracle():
    from oracle import bridge
    return bridge


def _workload(n_phi=None, spread_phi=None, spread_z=None, config=1):
    from trax_amd import events as ev
    w = ev.config_workload(config, 1)
    sc = w.seed_config
    sc.pair_generator = "sector"
    if n_phi is not None:
        sc.n_phi = n_phi
    for t in sc.triplets:
        if spread_phi is not None:
            t.pair_spread_phi = spread_phi
        if spread_z is not None:
            t.pair_spread_z = spread_z
    return w, sc, ev.make_batch(w)


def _cells64(sc, x, y, z):
    """(row, column) of each hit in float64, clamped like the grid build."""
    phi = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    zs = np.floor((z.astype(np.float64) + 300.0) / (600.0 / sc.n_z))
    ps = np.floor((phi + np.pi) / (2 * np.pi / sc.n_phi))
    return np.clip(zs, 0, sc.n_z).astype(np.int64), np.clip(ps, 0, sc.n_phi).astype(np.int64)


def test_sector_window_closed_form():
    B = _oracle()
    _, sc, _ = _workload()
    s = sc.triplets[0].pair_spread_phi
    sz = sc.triplets[0].pair_spread_z
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(400):
        phi = rng.uniform(-np.pi, np.pi)
        z = rng.uniform(-250, 250)
        x, y = np.float32(4.4 * np.cos(phi)), np.float32(4.4 * np.sin(phi))
        zs, ps = _cells64(sc, np.array([x]), np.array([y]), np.array([np.float32(z)]))
        q = (np.arctan2(np.float64(y), np.float64(x)) + np.pi) / (2 * np.pi / sc.n_phi)
        if abs(q - round(q)) < 1e-3 or abs((z + 300) / (600 / sc.n_z) % 1) < 1e-3:
            continue  # too close to a cell edge for a float64 reference
        zs, ps = int(zs[0]), int(ps[0])
        w = B.sector_window(sc, 0, float(x), float(y), float(np.float32(z)))
        assert w["zLowSector"] == (zs - sz) % 2**32
        assert w["zHighSector"] == min(zs + sz + 1, sc.n_z)
        lo, hi = ps - s, ps + s + 1
        wrap = lo < 0 or hi > sc.n_phi + 1
        assert w["wrap"] == wrap
        assert w["phiLowSector"] == (lo + sc.n_phi if lo < 0 else lo)
        assert w["phiHighSector"] == (hi - sc.n_phi if hi > sc.n_phi + 1 else hi)
        checked += 1
    assert checked > 300


def test_sector_window_underflow_below_row_spread():
    """zs < pairSpreadZ: zs - spread wraps as a uint, the row loop runs zero
    times (PairGeneratorSector.h:101-103)."""
    B = _oracle()
    _, sc, _ = _workload(spread_z=2)
    w = B.sector_window(sc, 0, 4.4, 0.1, -299.0)  # row 0
    assert w["zLowSector"] == 2**32 - 2 and w["zHighSector"] == 3


def _brute_force(sc, r, batch):
    """(segment, inner, outer) of every pair by cell membership."""
    E, L, P = batch.n_events, sc.n_layers, len(sc.triplets)
    C = sc.cells_per_layer
    zs, ps = _cells64(sc, r["x"], r["y"], r["z"])
    rows = []
    for e in range(E):
        for p, t in enumerate(sc.triplets):
            g1 = r["grid"][(e * L + t.layer1 - 1) * C:(e * L + t.layer1) * C]
            g2 = r["grid"][(e * L + t.layer2 - 1) * C:(e * L + t.layer2) * C]
            inner = np.arange(g1[0], g1[C - 1])
            outer = np.arange(g2[0], g2[C - 1])
            s, sz = t.pair_spread_phi, t.pair_spread_z
            for i in inner:
                z0, p0 = zs[i], ps[i]
                if z0 < sz:
                    continue
                rlo, rhi = z0 - sz, min(z0 + sz + 1, sc.n_z)
                lo, hi = p0 - s, p0 + s + 1
                if lo < 0 or hi > sc.n_phi + 1:  # wrap: [lo', nPhi) and [0, hi')
                    lo = lo + sc.n_phi if lo < 0 else lo
                    hi = hi - sc.n_phi if hi > sc.n_phi + 1 else hi
                    cols = set(range(lo, sc.n_phi)) | set(range(0, hi))
                else:
                    cols = set(range(lo, hi))
                sel = outer[(zs[outer] >= rlo) & (zs[outer] < rhi) & np.isin(ps[outer], list(cols))]
                rows.extend((e * P + p, i, j) for j in sel)
    return rows_sorted(np.array(rows, dtype=np.uint64).reshape(-1, 3))


@pytest.mark.parametrize("config", [0, 1])
def test_sector_pairs_equal_brute_force(config):
    B = _oracle()
    _, sc, batch = _workload(config=config)
    r = B.run_chain(sc, batch, stop_after=1)
    assert r["n_pairs_unwritten"] == 0
    got = rows_sorted(with_segment(written_pairs(r["pairs"]), r["pair_offsets"]))
    ref = _brute_force(sc, r, batch)
    assert len(ref) > 0
    assert np.array_equal(got, ref)


def test_sector_duplicate_visits_collapse():
    """A phi window of 2s+1 = 9 cells on an 8-cell ring visits one column
    twice: one pair in the bitmap, two visits in the counts."""
    B = _oracle()
    _, sc, batch = _workload(n_phi=8, spread_phi=4)
    r = B.run_chain(sc, batch, stop_after=1)
    assert r["n_pairs_unwritten"] > 0
    assert r["n_pairs_raw_overcount"] == r["n_pairs_unwritten"]
    got = rows_sorted(with_segment(written_pairs(r["pairs"]), r["pair_offsets"]))
    assert np.array_equal(got, _brute_force(sc, r, batch))
    # closed form of the repeated visits: the wrapped part of each row walks
    # [zStart, H) and the first part [A, zEnd); they meet on [A, min(H, zEnd))
    C = sc.cells_per_layer
    L = sc.n_layers
    extra = 0
    for p, t in enumerate(sc.triplets):
        g1 = r["grid"][(t.layer1 - 1) * C:t.layer1 * C]
        g2 = r["grid"][(t.layer2 - 1) * C:t.layer2 * C]
        for i in range(g1[0], g1[C - 1]):
            w = B.sector_window(sc, p, float(r["x"][i]), float(r["y"][i]), float(r["z"][i]))
            if not w["wrap"]:
                continue
            for zsec in range(w["zLowSector"], w["zHighSector"]):
                row = zsec * (sc.n_phi + 1)
                A, H = int(g2[row + w["phiLowSector"]]), int(g2[row + w["phiHighSector"]])
                zend = int(g2[row + sc.n_phi])
                extra += max(0, min(H, zend) - A)
    assert extra == r["n_pairs_raw_overcount"]
    assert L == 3


def test_sector_rejects_windows_beyond_the_row():
    """2 * pairSpreadPhi > nPhi would index outside the grid row (the
    reference reads neighbouring rows there); the oracle refuses it."""
    B = _oracle()
    _, sc, batch = _workload(n_phi=8, spread_phi=5)
    with pytest.raises(RuntimeError):
        B.run_chain(sc, batch, stop_after=1)
