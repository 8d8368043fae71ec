"""C ABI behaviour on the device: workspace checks, the segmented prefix
sum, the scan's look-back words, chain graphs and the plain method, against
numpy and the oracle."""
import ctypes as C

import numpy as np
import pytest

from helpers import compare_batch_per_event, oracle_per_event, rows_sorted, with_segment

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _oracle():
    from oracle import bridge
    return bridge


def _call_chain(ctx, cfg, bufs, db):
    from trax_amd import runtime
    hin, lay, hs, out = db.hits_c(), db.layout_c(), bufs.sorted_c(), bufs.outputs_c()
    return ctx.lib.trax_seed_chain(ctx.h, C.byref(cfg), C.byref(hin), C.byref(lay), C.byref(hs),
                                   C.c_void_p(bufs.grid.data_ptr()), C.byref(bufs.ws), C.byref(out),
                                   runtime.stream_ptr())


# ------------------------------------------------------------ workspace checks
def test_workspace_for_fewer_triplets_is_rejected(ctx):
    """A workspace laid out for one layer triplet, reused with a four-triplet
    configuration: the work offsets would overrun into the pair counts, so
    every stage call refuses it."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w4 = ev.config_workload(3, 2)
    batch = ev.make_batch(w4)
    sc1 = ev.config_workload(3, 2).seed_config
    sc1.triplets = sc1.triplets[:1]
    sc1.triplets[0].layer3 = 6  # same grid layers as the four-triplet table
    db = runtime.DeviceBatch.upload(batch)
    b1 = runtime.ChainBuffers(sc1, 2, batch.n_hits, 1 << 20, 1 << 21, 1 << 18, candidates=False)
    cfg4 = runtime.seed_config_c(w4.seed_config)
    rc = _call_chain(ctx, cfg4, b1, db)
    assert rc == -1
    assert b"fewer layer triplets" in ctx.lib.trax_ctx_last_error(ctx.h)
    # the same workspace with its own configuration runs
    assert _call_chain(ctx, b1.cfg, b1, db) == 0
    torch.cuda.synchronize()
    assert ctx.flags() == 0


def test_chain_rejects_more_hits_than_the_workspace(ctx):
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(1, 2)
    batch = ev.make_batch(w)
    db = runtime.DeviceBatch.upload(batch)
    small = runtime.ChainBuffers(w.seed_config, 2, batch.n_hits - 1, 1 << 20, 1 << 21, 1 << 18)
    assert _call_chain(ctx, small.cfg, small, db) == -1
    assert b"more hits" in ctx.lib.trax_ctx_last_error(ctx.h)


def test_shared_middle_layer_matches_oracle(ctx):
    """Layer triplets (1,3,4), (2,3,5), (3,4,6): two of them walk layer 3's
    hits as the middle layer, so the pair passes have up to twice as many
    work items as hits; staged and fused chains against the oracle."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(3, 2)
    w.seed_config.triplets = [ev.LayerTriplet(1, 3, 4), ev.LayerTriplet(2, 3, 5), ev.LayerTriplet(3, 4, 6)]
    sc, batch = w.seed_config, ev.make_batch(w)
    bufs, _ = runtime.run_batch(ctx, sc, batch)
    g = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    fb, _ = runtime.run_batch(ctx, sc, batch, candidates=False)
    f = runtime.fetch(fb, batch.n_events, batch.n_hits)
    o = _oracle().run_chain(sc, batch)
    assert bufs.ws.pair_items == 2 * 2 * bufs.max_hits
    assert np.array_equal(g["pair_offsets"], o["pair_offsets"])
    assert np.array_equal(g["tracklet_offsets"], o["tracklet_offsets"])
    assert np.array_equal(rows_sorted(with_segment(g["tracklets"], g["tracklet_offsets"])),
                          rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"])))
    for k in ("pairs", "pair_offsets", "cand_offsets", "tracklets", "tracklet_offsets"):
        assert np.array_equal(f[k], g[k]), k
    assert len(o["tracklets"]) > 100


# ------------------------------------------------------------ segmented scan
def _seg_want(a, off):
    out = np.empty_like(a)
    for k in range(len(off) - 1):
        lo, hi = int(off[k]), int(off[k + 1])
        if hi > lo:
            out[lo] = 0
            np.cumsum(a[lo:hi - 1], dtype=np.uint32, out=out[lo + 1:hi])
    return out


@pytest.mark.parametrize("n,n_seg,empty", [(1, 1, False), (1000, 1, False), (1000, 1000, False),
                                           (100_000, 37, True), (3_000_001, 5000, True),
                                           (1 << 21, 64, False)])
def test_prefix_sum_segmented(ctx, n, n_seg, empty):
    from trax_amd import runtime
    rng = np.random.default_rng(n + n_seg)
    a = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    cuts = np.sort(rng.integers(0, n + 1, n_seg - 1)) if n_seg > 1 else np.zeros(0, np.int64)
    if n_seg == n:
        cuts = np.arange(1, n)
    if empty:  # repeated cuts: empty segments inside, and an empty one at each end
        cuts = np.concatenate([cuts, cuts[: n_seg // 4], [0, n]])
    off = np.concatenate([[0], np.sort(cuts), [n]]).astype(np.uint32)
    print(f"segmented scan n={n} segments={len(off) - 1}", flush=True)
    t = torch.from_numpy(a.view(np.int32).copy()).cuda()
    d_off = torch.from_numpy(off.view(np.int32)).cuda()
    runtime.prefix_sum_segmented(ctx, t, d_off)
    got = t.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, _seg_want(a, off))


def test_prefix_sum_segmented_after_scans(ctx):
    """Segment bases are parked in the look-back words: ordinary scans before
    and after must be unaffected."""
    from trax_amd import runtime
    rng = np.random.default_rng(5)
    for n in (300_000, 2_500_000, 40_000):
        a = rng.integers(0, 1000, n, dtype=np.uint32)
        off = np.concatenate([[0], np.sort(rng.integers(0, n, 99)), [n]]).astype(np.uint32)
        t = torch.from_numpy(a.view(np.int32).copy()).cuda()
        runtime.prefix_sum_segmented(ctx, t, torch.from_numpy(off.view(np.int32)).cuda())
        assert np.array_equal(t.cpu().numpy().view(np.uint32), _seg_want(a, off))
        t2 = torch.from_numpy(a.view(np.int32).copy()).cuda()
        runtime.prefix_sum(ctx, t2)
        assert np.array_equal(t2.cpu().numpy().view(np.uint32), _oracle().scan_u32(a))


def test_scan_partials_never_look_like_published_tiles():
    """The large-array form parks its chunk partials in the look-back words.
    Partials whose upper 32 bits read as (epoch << 2) | 2 for epochs
    0..63 (chunk sums 2, 4, 0, 4, 0, ...) must not be taken as published
    tiles by the single-pass scans of a fresh context (epochs 0, 1, 2, ...)."""
    from trax_amd import runtime
    c = runtime.Context(0)
    try:
        n_big = 1 << 21  # 128 chunks of 16384
        sums = np.zeros(128, np.uint32)
        sums[0] = 2
        sums[1::2] = 4
        big = np.zeros(n_big, np.uint32)
        big[::16384] = sums
        t = torch.from_numpy(big.view(np.int32).copy()).cuda()
        runtime.prefix_sum(c, t)
        assert np.array_equal(t.cpu().numpy().view(np.uint32), _oracle().scan_u32(big))
        rng = np.random.default_rng(9)
        for e in range(64):
            a = rng.integers(0, 1 << 32, 600_000, dtype=np.uint64).astype(np.uint32)
            s = torch.from_numpy(a.view(np.int32).copy()).cuda()
            runtime.prefix_sum(c, s)
            assert np.array_equal(s.cpu().numpy().view(np.uint32), _oracle().scan_u32(a)), e
    finally:
        c.close()


# ---------------------------------------------------------------- chain graph
def test_chain_graph_replays_match_eager_and_oracle(ctx):
    """trax_chain_graph_capture records the fused chain once; three
    launches, with unrelated eager work and host synchronisation between
    them, each give the eager chain's arrays, which match the oracle."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(3, 3)
    sc, batch = w.seed_config, ev.make_batch(w)
    bufs, db = runtime.run_batch(ctx, sc, batch, candidates=False)
    eager = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    keys = ("grid", "orig_index", "pairs", "pair_offsets", "cand_offsets", "tracklets", "tracklet_offsets")
    g = runtime.ChainGraph(ctx, bufs, db)
    try:
        for rep in range(3):
            for k in ("pairs", "tracklet_offsets"):  # scribble over the outputs
                getattr(bufs, k).fill_(-1)
            for t in bufs.t:
                t.fill_(-1)
            bufs.grid.fill_(-1)
            side = torch.randn(1 << 20, device="cuda").cumsum(0)
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            assert ctx.flags() == 0
            got = runtime.fetch(bufs, batch.n_events, batch.n_hits)
            for k in keys:
                assert np.array_equal(got[k], eager[k]), (rep, k)
            del side
    finally:
        g.close()
    o = _oracle().run_chain(sc, batch)
    assert np.array_equal(eager["tracklet_offsets"], o["tracklet_offsets"])
    assert np.array_equal(rows_sorted(with_segment(eager["tracklets"], eager["tracklet_offsets"])),
                          rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"])))


def test_chain_graph_capture_rejects_bad_arguments(ctx):
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(1, 1)
    batch = ev.make_batch(w)
    db = runtime.DeviceBatch.upload(batch)
    small = runtime.ChainBuffers(w.seed_config, 1, batch.n_hits - 1, 1 << 20, 1 << 21, 1 << 18)
    hin, lay, hs, out = db.hits_c(), db.layout_c(), small.sorted_c(), small.outputs_c()
    h = C.c_void_p()
    rc = ctx.lib.trax_chain_graph_capture(ctx.h, C.byref(small.cfg), C.byref(hin), C.byref(lay), C.byref(hs),
                                          C.c_void_p(small.grid.data_ptr()), C.byref(small.ws), C.byref(out),
                                          C.byref(h))
    assert rc == -1 and not h.value
    assert ctx.lib.trax_chain_graph_launch(None, None) == -1
    assert ctx.lib.trax_chain_graph_destroy(None) == 0
    # the context still works after the refused capture
    bufs, _ = runtime.run_batch(ctx, w.seed_config, batch)
    assert runtime.totals(bufs)[2] > 0


# ---------------------------------------------------------------- plain path
@pytest.mark.parametrize("config_id", [0, 1])
def test_plain_method_matches_oracle(ctx, config_id):
    """The plain method (one event, two-phase stages with a blocking host read
    of each total) against the oracle."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(config_id, 1)
    sc, batch = w.seed_config, ev.make_batch(w)
    db = runtime.DeviceBatch.upload(batch)
    p, c, t = runtime.default_capacities(batch.n_hits)
    bufs = runtime.ChainBuffers(sc, 1, batch.n_hits, p, c, t)
    n = runtime.run_plain_event(ctx, bufs, db)
    assert ctx.flags() == 0
    g = runtime.fetch(bufs, 1, batch.n_hits)
    o = _oracle().run_chain(sc, batch)
    assert n == len(o["tracklets"]) > 0
    compare_batch_per_event(g, batch, len(sc.triplets), [o])
