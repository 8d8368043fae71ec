"""GPU parity of the grid build alone (GridBuilder.h:48-427) on hand-made
layouts that reach every path of the sort: segments sorted in LDS, segments
of more than 8192 hits and grids of more than 2^19 cells sorted in place in
the output columns, empty segments, every hit of a segment in one cell
(stability), and hits in the padding row and column (z >= 300, phi == pi).
The grid (cell starts) and every sorted column must equal the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _batch(layers_per_event, n_layers, rng):
    """EventBatch from per-event lists of per-layer (x, y, z) arrays."""
    from trax_amd import events as ev
    cols = {k: [] for k in ("x", "y", "z", "layer", "det_id", "hit_id", "evt")}
    ev_off, ln, lo = [], [], []
    n = 0
    for e, layers in enumerate(layers_per_event):
        ev_off.append(n)
        rel = 0
        for li in range(n_layers):
            x, y, z = layers[li]
            k = len(x)
            cols["x"].append(np.asarray(x, np.float32))
            cols["y"].append(np.asarray(y, np.float32))
            cols["z"].append(np.asarray(z, np.float32))
            cols["layer"].append(np.full(k, li + 1, np.uint32))
            cols["det_id"].append(rng.integers(0, 1 << 30, k).astype(np.uint32))
            cols["hit_id"].append(np.arange(n + rel, n + rel + k, dtype=np.uint32))
            cols["evt"].append(np.full(k, e, np.uint32))
            ln.append(k)
            lo.append(rel)
            rel += k
        n += rel
    cat = {k: np.concatenate(v) if v else np.zeros(0) for k, v in cols.items()}
    return ev.EventBatch(cat["x"], cat["y"], cat["z"], cat["layer"], cat["det_id"], cat["hit_id"],
                         cat["evt"], np.asarray(ev_off, np.uint32), np.asarray(ln, np.uint32),
                         np.asarray(lo, np.uint32), n_layers)


def _cylinder(rng, n, r, zlo=-30.0, zhi=30.0):
    phi = rng.uniform(-np.pi, np.pi, n)
    return r * np.cos(phi), r * np.sin(phi), rng.uniform(zlo, zhi, n)


def _grid_parity(ctx, sc, batch):
    from oracle import bridge
    from trax_amd import runtime
    print(f"grid parity: {batch.n_events} events, {batch.n_hits} hits, "
          f"segments {list(batch.layer_nhits)}, grid {sc.n_phi}x{sc.n_z}", flush=True)
    db = runtime.DeviceBatch.upload(batch)
    bufs = runtime.ChainBuffers(sc, batch.n_events, max(batch.n_hits, 1), 1024, 2048, 1024)
    ctx.flags(clear=True)
    runtime.grid_build(ctx, bufs, db)
    torch.cuda.synchronize()
    assert ctx.flags() == 0
    o = bridge.run_chain(sc, batch, stop_after=0)
    n = batch.n_hits
    G = batch.n_events * sc.n_layers * sc.cells_per_layer
    grid = bufs.grid[:G].cpu().numpy().view(np.uint32)
    assert np.array_equal(grid, o["grid"])
    s = bufs.sorted
    for k in ("x", "y", "z"):
        assert np.array_equal(s[k][:n].cpu().numpy().view(np.uint32), o[k].view(np.uint32)), k
    for k in ("layer", "det_id", "hit_id", "evt", "orig_index"):
        assert np.array_equal(s[k][:n].cpu().numpy().view(np.uint32), o[k]), k
    phi = bridge.eval_f32("atan2", o["y"], o["x"])
    assert np.array_equal(s["phi"][:n].cpu().numpy().view(np.uint32), phi.view(np.uint32))
    zphi = s["zphi"][:2 * n].cpu().numpy().reshape(-1, 2)
    assert np.array_equal(zphi[:, 0].view(np.uint32), o["z"].view(np.uint32))
    assert np.array_equal(zphi[:, 1].view(np.uint32), phi.view(np.uint32))
    xyzphi = s["xyzphi"][:4 * n].cpu().numpy().reshape(-1, 4)
    for c, want in enumerate((o["x"], o["y"], o["z"], phi)):
        assert np.array_equal(xyzphi[:, c].view(np.uint32), want.view(np.uint32)), ("xyzphi", c)
    return o


def _config(n_z, n_phi):
    from trax_amd import events as ev
    sc = ev.config_workload(1, 1).seed_config
    sc.n_z, sc.n_phi = n_z, n_phi
    return sc


def test_grid_large_and_empty_segments(ctx):
    """Layer 1: 20000 hits (in-place sort), layer 2: empty, layer 3: 8192 hits
    (the largest LDS segment); second event: 8193, 1, 0 hits."""
    rng = np.random.default_rng(5)
    e0 = [_cylinder(rng, 20000, 4.4), _cylinder(rng, 0, 7.3), _cylinder(rng, 8192, 10.2)]
    e1 = [_cylinder(rng, 8193, 4.4), _cylinder(rng, 1, 7.3), _cylinder(rng, 0, 10.2)]
    _grid_parity(ctx, _config(64, 128), _batch([e0, e1], 3, rng))


def test_grid_more_than_2_19_cells(ctx):
    """1024 x 1024 cells (20-bit cell ids): every segment takes the in-place
    sort with three digit passes."""
    rng = np.random.default_rng(6)
    ev_ = [[_cylinder(rng, 3000, 4.4), _cylinder(rng, 2500, 7.3), _cylinder(rng, 50, 10.2)]
           for _ in range(2)]
    _grid_parity(ctx, _config(1023, 1023), _batch(ev_, 3, rng))


def test_grid_one_cell_keeps_input_order(ctx):
    """5000 identical hits in one cell, plus a wide segment: the sorted order
    inside a cell is the input order."""
    rng = np.random.default_rng(7)
    same = (np.full(5000, 3.0), np.full(5000, 3.2), np.full(5000, 1.5))
    e0 = [same, _cylinder(rng, 4000, 7.3, -290, 290), _cylinder(rng, 100, 10.2)]
    o = _grid_parity(ctx, _config(256, 128), _batch([e0], 3, rng))
    assert np.array_equal(o["orig_index"][:5000], np.arange(5000))


def test_grid_padding_row_and_column(ctx):
    """z beyond +-300 cm (clamped to the padding row / row 0) and phi == +-pi
    (atan2(+-0, -1): padding column / column 0), mixed with ordinary hits."""
    rng = np.random.default_rng(8)
    x = np.concatenate([[-4.4, -4.4, -4.4, -4.4, 4.4, 4.4], _cylinder(rng, 500, 4.4)[0]])
    y = np.concatenate([[0.0, -0.0, 0.0, -0.0, 0.0, 0.0], _cylinder(rng, 500, 4.4)[1]])
    z = np.concatenate([[301.0, -301.0, 300.0, 0.0, 299.9, -300.0],
                        rng.uniform(-400, 400, 500)])
    layers = [(x, y, z), _cylinder(rng, 300, 7.3), _cylinder(rng, 300, 10.2)]
    _grid_parity(ctx, _config(16, 32), _batch([layers, layers], 3, rng))


@pytest.mark.parametrize("waves", [0, 1, 2, 4, 8, 16])
def test_grid_block_shapes(ctx, waves):
    """Every grid-sort block shape (0: chosen from the mean segment size)
    gives the same grid on a PU200 event."""
    from trax_amd import events as ev
    w = ev.config_workload(3, 2)
    try:
        ctx.set_launch(grid_waves=waves)
        _grid_parity(ctx, w.seed_config, ev.make_batch(w))
    finally:
        ctx.set_launch()


@pytest.mark.parametrize("waves", [0, 1, 8])
def test_many_small_segments_with_outliers(ctx, waves):
    """200 events of a few hits per layer make the sort's LDS cap small
    (1024 slots); a 900-hit segment then sorts by radix in LDS and a
    6000-hit segment in place in the output columns, next to the tiny
    ones, under every block shape."""
    rng = np.random.default_rng(12)
    events = []
    for e in range(200):
        layers = [_cylinder(rng, int(rng.integers(0, 9)), r) for r in (4.4, 7.3, 10.2)]
        if e == 3:
            layers[1] = _cylinder(rng, 900, 7.3)
        if e == 150:
            layers[0] = _cylinder(rng, 6000, 4.4)
        events.append(layers)
    try:
        ctx.set_launch(grid_waves=waves)
        _grid_parity(ctx, _config(16, 32), _batch(events, 3, rng))
    finally:
        ctx.set_launch()
