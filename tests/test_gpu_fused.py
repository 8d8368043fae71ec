"""GPU parity of the fused prediction + filter pass
(trax_triplet_predict_filter): the same triplets, in the same order, and the
same triplet and candidate offsets and totals as the staged chain
(TripletThetaPhiPredictor.h:43-44 then TripletThetaPhiFilter.h:46-48), and
the same triplet multisets per segment as the oracle.  Covers the common
path (tiles whose candidates fit the lane buffers and whose accepted records
fit the slab), tiles with a pair past the lane buffer (the rest in the wave's
spill pool; tiles that overflow the pool, redone by a second kernel),
tiles whose accepted records overflow the slab (walked again at placement),
triplet-capacity overflow and regrowth, and graph replay."""
import numpy as np
import pytest

from helpers import rows_sorted, with_segment

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _oracle():
    from oracle import bridge
    return bridge


def _both(ctx, sc, batch):
    from trax_amd import runtime
    print(f"fused vs staged: {batch.n_events} events, {batch.n_hits} hits", flush=True)
    bs, _ = runtime.run_batch(ctx, sc, batch)
    bf, df = runtime.run_batch(ctx, sc, batch, candidates=False)
    torch.cuda.synchronize()
    assert bf.cands is None
    s = runtime.fetch(bs, batch.n_events, batch.n_hits)
    f = runtime.fetch(bf, batch.n_events, batch.n_hits)
    assert runtime.totals(bs) == runtime.totals(bf)
    for k in ("grid", "pairs", "pair_offsets", "cand_offsets", "tracklets", "tracklet_offsets"):
        assert np.array_equal(s[k], f[k]), k
    return s, f, (bf, df)


FUSE_BUF, SPILL = 8, 192  # lane buffer depth and spill pool size of the fused walk


def _spill_per_tile(per_pair):
    """Filled slots past the lane buffers, per tile of 64 pairs (a lower
    bound of the tile's spill pool use)."""
    ntile = (len(per_pair) + 63) // 64
    extra = np.zeros(ntile * 64, np.int64)
    extra[:len(per_pair)] = np.maximum(per_pair - FUSE_BUF, 0)
    return extra.reshape(ntile, 64).sum(1)


def _accepted_per_tile(s):
    """Accepted candidates per tile of 64 pairs and filled slots per pair,
    from the staged chain's candidates (triplets are the accepted candidates
    in candidate order)."""
    from trax_amd import runtime
    pairs, cands, trk = s["pairs"], s["cands"], s["tracklets"]
    filled = cands[:, 0] != runtime.UNWRITTEN
    per_pair = np.bincount(cands[filled, 0].astype(np.int64), minlength=len(pairs))
    acc = np.zeros(len(pairs), np.int64)
    j = 0
    for q, h3 in cands[filled]:
        if j < len(trk) and trk[j, 0] == pairs[q, 0] and trk[j, 1] == pairs[q, 1] and trk[j, 2] == h3:
            acc[q] += 1
            j += 1
    assert j == len(trk)
    ntile = (len(pairs) + 63) // 64
    per_tile = np.zeros(ntile * 64, np.int64)
    per_tile[:len(acc)] = acc
    return per_pair, per_tile.reshape(ntile, 64).sum(1)


CASES = [(0, 1), (1, 1), (1, 3), (2, 4), (3, 2)]


@pytest.mark.parametrize("config_id,n_events", CASES)
def test_fused_equals_staged_and_oracle(ctx, config_id, n_events):
    from trax_amd import events as ev
    w = ev.config_workload(config_id, n_events)
    batch = ev.make_batch(w)
    s, f, _ = _both(ctx, w.seed_config, batch)
    o = _oracle().run_chain(w.seed_config, batch)
    assert np.array_equal(f["tracklet_offsets"], o["tracklet_offsets"])
    gt = rows_sorted(with_segment(f["tracklets"], f["tracklet_offsets"]))
    ot = rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"]))
    assert np.array_equal(gt, ot)
    assert len(ot) > 0


def test_fused_wide_windows_spill(ctx):
    """Wide prediction windows: pairs with more candidate hits than the lane
    buffer spill into the wave's pool."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(1, 1)
    for t in w.seed_config.triplets:
        t.sigma_phi, t.sigma_z = 3.0, 1.0
    batch = ev.make_batch(w)
    s, f, _ = _both(ctx, w.seed_config, batch)
    per_pair, _ = _accepted_per_tile(s)
    assert per_pair.max() > runtime.PREDICT_BUFFER
    spill = _spill_per_tile(per_pair)
    assert ((spill > 0) & (spill <= SPILL // 2)).any()  # the spill pool in use
    assert len(f["tracklets"]) > 0


def test_fused_slab_overflow_tiles(ctx):
    """Wide windows and cuts that accept nearly everything: tiles with more
    accepted records than the slab holds are walked again at placement."""
    from trax_amd import events as ev
    w = ev.config_workload(1, 1)
    for t in w.seed_config.triplets:
        t.sigma_phi, t.sigma_z = 3.0, 1.0
        t.dphi_cut, t.dtheta_cut, t.tip_cut = 10.0, 100.0, 1.0e4
    batch = ev.make_batch(w)
    s, f, _ = _both(ctx, w.seed_config, batch)
    per_pair, per_tile = _accepted_per_tile(s)
    assert per_tile.max() > 256  # the slab holds 256 records per tile
    assert _spill_per_tile(per_pair).max() > SPILL  # pool overflow: the tile is redone
    o = _oracle().run_chain(w.seed_config, batch)
    gt = rows_sorted(with_segment(f["tracklets"], f["tracklet_offsets"]))
    ot = rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"]))
    assert np.array_equal(gt, ot)


def test_fused_tracklet_overflow_then_regrow(ctx):
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(1, 2)
    batch = ev.make_batch(w)
    s, _, _ = _both(ctx, w.seed_config, batch)
    n_t = len(s["tracklets"])
    assert n_t > 10
    db = runtime.DeviceBatch.upload(batch)
    small = runtime.ChainBuffers(w.seed_config, db.n_events, db.n_hits, 64 * db.n_hits, 0, n_t // 3,
                                 candidates=False)
    ctx.flags(clear=True)
    runtime.seed_chain(ctx, small, db)
    torch.cuda.synchronize()
    assert ctx.flags(clear=True) == runtime.FLAG_TRACKLET_OVERFLOW
    assert runtime.totals(small)[2] == n_t
    bufs, _ = runtime.run_batch(ctx, w.seed_config, batch, bufs=small, candidates=False)
    torch.cuda.synchronize()
    f = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    assert np.array_equal(f["tracklets"], s["tracklets"])


def test_fused_graph_replays(ctx):
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(1, 3)
    batch = ev.make_batch(w)
    s, _, (bf, df) = _both(ctx, w.seed_config, batch)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        runtime.seed_chain(ctx, bf, df)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        runtime.seed_chain(ctx, bf, df)
    for _ in range(3):
        bf.t[0].fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        f = runtime.fetch(bf, batch.n_events, batch.n_hits)
        assert np.array_equal(f["tracklets"], s["tracklets"])
        assert np.array_equal(f["tracklet_offsets"], s["tracklet_offsets"])
    assert ctx.flags() == 0
