"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Integer and index outputs must match exactly; the FP32 cut arithmetic is
pinned bit for bit by running the same FP32 definitions on both sides.
Pairs, candidates and triplets are compared as multisets per
(event, layer triplet) segment, because the reference's own output order
depends on its work-group size (PairGeneratorBeamspot.h:488-516); the
per-segment offsets must be identical arrays.
"""
import numpy as np
import pytest

from helpers import cand_rows, rows_sorted, with_segment, written_pairs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _oracle():
    from oracle import bridge
    return bridge


# ------------------------------------------------------------------ FP32 math
@pytest.mark.parametrize("op", ["atan2", "atan", "acos", "sign", "div", "sqrt"])
def test_fp32_definitions_bitwise(ctx, op):
    from trax_amd import runtime
    rng = np.random.default_rng(7)
    n = 1 << 20
    if op == "acos":
        a = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    elif op == "sqrt":
        a = np.abs(rng.standard_normal(n).astype(np.float32)) * 100
    else:
        a = (rng.standard_normal(n) * rng.choice([1e-3, 1.0, 50.0, 1e4], n)).astype(np.float32)
    b = (rng.standard_normal(n) * rng.choice([1e-3, 1.0, 50.0, 1e4], n)).astype(np.float32)
    special = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45,
                        3.4e38, -3.4e38, 0.5, -0.5], np.float32)
    if op in ("atan2", "div"):
        sa, sb = np.meshgrid(special, special)
        a = np.concatenate([a, sa.ravel()])
        b = np.concatenate([b, sb.ravel()])
    else:
        a = np.concatenate([a, special])
        b = np.concatenate([b, special])
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    got = runtime.fp32_eval(ctx, op, da, db).cpu().numpy()
    want = _oracle().eval_f32(op, a, b)
    gb, wb = got.view(np.uint32), want.view(np.uint32)
    nan_both = np.isnan(got) & np.isnan(want)
    bad = (gb != wb) & ~nan_both
    assert not bad.any(), f"{op}: {bad.sum()} mismatches, first at {np.argmax(bad)}: a={a[bad][:3]} b={b[bad][:3]} got={got[bad][:3]} want={want[bad][:3]}"


# ----------------------------------------------------------------- prefix sum
@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 5, 1023, 1024, 1025, 8191, 8192, 8193, 65537,
                               (1 << 20) - 1, 1 << 20, (1 << 21) - 1, 1 << 21, (1 << 21) + 16385,
                               (1 << 24) + 1])
def test_prefix_sum_sizes(ctx, n):
    from trax_amd import runtime
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1 << 32, max(n, 1), dtype=np.uint64).astype(np.uint32)[:n]
    t = torch.from_numpy(a.view(np.int32).copy()).cuda()
    runtime.prefix_sum(ctx, t)
    got = t.cpu().numpy().view(np.uint32)
    want = _oracle().scan_u32(a)
    assert np.array_equal(got, want)


def test_prefix_sum_known_answer(ctx):
    """trax_test.cpp:437-479: all ones, N = 2^20, last element N - 1."""
    from trax_amd import runtime
    n = 1 << 20
    t = torch.ones(n, dtype=torch.int32, device="cuda")
    runtime.prefix_sum(ctx, t)
    got = t.cpu().numpy()
    assert got[-1] == n - 1
    assert np.array_equal(got, np.arange(n, dtype=np.int32))


def test_prefix_sum_large_wraps(ctx):
    """10^8 elements, values near 2^32: mod-2^32 wrap and many look-back tiles."""
    from trax_amd import runtime
    n = 100_000_000
    g = torch.Generator(device="cuda").manual_seed(3)
    t = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    host = t.cpu().numpy().view(np.uint32).copy()
    runtime.prefix_sum(ctx, t)
    got = t.cpu().numpy().view(np.uint32)
    want = np.empty_like(host)
    want[0] = 0
    np.cumsum(host[:-1], dtype=np.uint32, out=want[1:])
    assert np.array_equal(got, want)


def test_prefix_sum_repeated_launches(ctx):
    """Status words are epoch-tagged: back-to-back launches must not see stale tiles."""
    from trax_amd import runtime
    rng = np.random.default_rng(11)
    for n in [300_000, 5_000, 2_000_000, 17, 1_000_000]:
        a = rng.integers(0, 1000, n, dtype=np.uint32)
        t = torch.from_numpy(a.view(np.int32).copy()).cuda()
        runtime.prefix_sum(ctx, t)
        assert np.array_equal(t.cpu().numpy().view(np.uint32), _oracle().scan_u32(a))


# -------------------------------------------------------------------- chain
CASES = [(0, 1), (1, 1), (1, 3), (2, 4), (3, 2)]


def _compare_chain(ctx, sc, batch):
    from trax_amd import runtime
    print(f"chain parity: {batch.n_events} events, {batch.n_hits} hits", flush=True)
    bufs, db = runtime.run_batch(ctx, sc, batch)
    torch.cuda.synchronize()
    g = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    o = _oracle().run_chain(sc, batch)
    # grid stage: identical cell starts, identical (stable) sorted columns
    assert np.array_equal(g["grid"], o["grid"])
    for k in ("x", "y", "z"):
        assert np.array_equal(g[k].view(np.uint32), o[k].view(np.uint32)), k
    for k in ("layer", "det_id", "hit_id", "evt", "orig_index"):
        assert np.array_equal(g[k], o[k]), k
    phi_ref = _oracle().eval_f32("atan2", o["y"], o["x"])
    assert np.array_equal(g["phi"].view(np.uint32), phi_ref.view(np.uint32))
    # pairs: identical per-segment offsets and pair multisets
    assert o["n_pairs_unwritten"] == 0
    assert np.array_equal(g["pair_offsets"], o["pair_offsets"])
    gp = rows_sorted(with_segment(g["pairs"], g["pair_offsets"]))
    op_ = rows_sorted(with_segment(written_pairs(o["pairs"]), o["pair_offsets"]))
    assert np.array_equal(gp, op_)
    # candidates: same (hit1, hit2, hit3) multisets per segment
    assert np.array_equal(g["cand_offsets"], o["cand_offsets"])
    assert (g["cands"][:, 0] == 0xFFFFFFFF).sum() == o["n_cands_unwritten"]
    gc = rows_sorted(cand_rows(g["pairs"], g["cands"], g["cand_offsets"]))
    oc = rows_sorted(cand_rows(o["pairs"], o["cands"], o["cand_offsets"]))
    assert np.array_equal(gc, oc)
    # triplets
    assert np.array_equal(g["tracklet_offsets"], o["tracklet_offsets"])
    gt = rows_sorted(with_segment(g["tracklets"], g["tracklet_offsets"]))
    ot = rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"]))
    assert np.array_equal(gt, ot)
    return g, o


@pytest.mark.parametrize("config_id,n_events", CASES)
def test_chain_matches_oracle(ctx, config_id, n_events):
    from trax_amd import events as ev
    w = ev.config_workload(config_id, n_events)
    batch = ev.make_batch(w)
    g, o = _compare_chain(ctx, w.seed_config, batch)
    assert len(o["tracklets"]) > 0


def test_chain_reference_cut_row(ctx):
    """The reference's own layer-triplet row (sigmaZ != sigmaPhi, so predictor
    count and store windows differ) on pile-up events."""
    from trax_amd import events as ev
    w = ev.config_workload(1, 2)
    t = w.seed_config.triplets[0]
    t.sigma_z, t.sigma_phi, t.dphi_cut, t.z0 = 0.037, 0.032, 0.075, 10
    _compare_chain(ctx, w.seed_config, ev.make_batch(w))


def test_batched_equals_single_events(ctx):
    """An event's triplets are the same alone and inside a batch."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(2, 5)
    batch = ev.make_batch(w)
    bufs, _ = runtime.run_batch(ctx, w.seed_config, batch)
    g = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    P = len(w.seed_config.triplets)
    for e in range(batch.n_events):
        single = ev.make_batch(w, events=[e])
        b1, _ = runtime.run_batch(ctx, w.seed_config, single)
        s = runtime.fetch(b1, 1, single.n_hits)
        base = int(batch.event_offset[e])
        lo, hi = g["tracklet_offsets"][e * P], g["tracklet_offsets"][(e + 1) * P]
        mine = g["orig_index"][g["tracklets"][lo:hi]] - base
        alone = s["orig_index"][s["tracklets"]]
        assert np.array_equal(rows_sorted(mine), rows_sorted(alone))


def test_single_call_predictor_matches_two_phase(ctx):
    """The tiled predictor (one walk, scan, placement) writes the same
    candidate slots as count + scan + store, including pairs whose budget and
    store list are longer than the lane buffer (they walk again)."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(3, 2)
    batch = ev.make_batch(w)
    bufs, db = runtime.run_batch(ctx, w.seed_config, batch)
    g = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    written = g["cands"][:, 0] != 0xFFFFFFFF
    per_pair = np.bincount(g["cands"][written, 0])
    assert per_pair.max() > runtime.PREDICT_BUFFER  # the re-walk path runs
    runtime.predict_count(ctx, bufs, db.n_events)
    runtime.predict_store(ctx, bufs, db.n_events)
    torch.cuda.synchronize()
    t = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    assert np.array_equal(g["cands"], t["cands"])
    assert np.array_equal(g["cand_offsets"], t["cand_offsets"])
    assert ctx.flags() == 0


@pytest.mark.parametrize("sigma_phi,sigma_z", [(1.5, 0.1), (3.0, 1.0)])
def test_chain_wide_windows(ctx, sigma_phi, sigma_z):
    """Wide prediction windows: pairs with more store hits than the lane
    buffer and tiles whose slots overflow their stash slab both walk again;
    the chain must still match the oracle, and the single-call predictor the
    two-phase one."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(1, 1)
    for t in w.seed_config.triplets:
        t.sigma_phi, t.sigma_z = sigma_phi, sigma_z
    batch = ev.make_batch(w)
    g, o = _compare_chain(ctx, w.seed_config, batch)
    co = g["cand_offsets"]
    written = g["cands"][:, 0] != 0xFFFFFFFF
    assert np.bincount(g["cands"][written, 0]).max() > runtime.PREDICT_BUFFER
    bufs, db = runtime.run_batch(ctx, w.seed_config, batch)
    runtime.predict_count(ctx, bufs, db.n_events)
    runtime.predict_store(ctx, bufs, db.n_events)
    torch.cuda.synchronize()
    t = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    assert np.array_equal(g["cands"], t["cands"])
    assert np.array_equal(co, t["cand_offsets"])


def test_pair_store_past_the_count_masks(ctx):
    """Coarse bins give pair windows of hundreds of visited hits.  The chain's
    pair store takes the count pass's verdicts on each work item's first 64
    visited hits and tests the rest itself; the separate count and store
    entry points pass no verdicts.  Both must give the oracle's pairs."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(3, 1)
    w.seed_config.n_z, w.seed_config.n_phi = 16, 16
    batch = ev.make_batch(w)
    g, o = _compare_chain(ctx, w.seed_config, batch)
    # an outer hit with more than 2 x 64 pairs has a work item past its mask
    assert np.bincount(g["pairs"][:, 1]).max() > 128
    bufs, db = runtime.run_batch(ctx, w.seed_config, batch)
    runtime.pairs_count(ctx, bufs, db.n_events)
    runtime.pairs_store(ctx, bufs, db.n_events)
    torch.cuda.synchronize()
    t = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    assert np.array_equal(g["pair_offsets"], t["pair_offsets"])
    assert np.array_equal(g["pairs"], t["pairs"])
    assert ctx.flags() == 0


def _event_without_layers(b1, layers):
    """A one-event batch with the hits of `layers` (1-based) removed."""
    from trax_amd import events as ev
    keep = ~np.isin(b1.layer, np.asarray(layers, np.uint32))
    nh = b1.layer_nhits.copy()
    for l in layers:
        nh[l - 1] = 0
    off = np.concatenate([[0], np.cumsum(nh)[:-1]]).astype(np.uint32)
    cols = [np.ascontiguousarray(c[keep]) for c in (b1.x, b1.y, b1.z, b1.layer, b1.det_id, b1.hit_id, b1.evt)]
    return ev.EventBatch(*cols, np.zeros(1, np.uint32), nh.astype(np.uint32), off, b1.n_layers)


def _empty_event(L):
    from trax_amd import events as ev
    return ev.EventBatch(*[np.zeros(0, np.float32)] * 3, *[np.zeros(0, np.uint32)] * 4, np.zeros(1, np.uint32),
                         np.zeros(L, np.uint32), np.zeros(L, np.uint32), L)


@pytest.mark.parametrize("pair_generator", ["beamspot", "sector"])
def test_chain_ragged_and_empty_events(ctx, pair_generator):
    """A batch of uneven events: an empty event first and last, events with
    their first, middle or outer layer empty, and an event with hits on one
    layer only.  Every stage matches the oracle (segments of no pairs and
    no triplets included), with either pair generator, and the fused pass
    gives the staged triplets."""
    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(2, 4)
    w.seed_config.pair_generator = pair_generator
    full = ev.make_batch(w)
    L = w.seed_config.n_layers
    parts = [_empty_event(L), full.event(0), _event_without_layers(full.event(1), [1]),
             _event_without_layers(full.event(2), [2]), _empty_event(L),
             _event_without_layers(full.event(3), [4]), _event_without_layers(full.event(0), [2, 3, 4]),
             full.event(1), _empty_event(L)]
    batch = ev.concat_events(parts)
    g, o = _compare_chain(ctx, w.seed_config, batch)
    P = len(w.seed_config.triplets)
    toff = g["tracklet_offsets"]
    assert toff[P] == 0 and toff[-1] - toff[-1 - P] == 0  # the empty events hold no triplets
    assert len(o["tracklets"]) > 0
    fbufs, _ = runtime.run_batch(ctx, w.seed_config, batch, candidates=False)
    torch.cuda.synchronize()
    f = runtime.fetch(fbufs, batch.n_events, batch.n_hits)
    for k in ("tracklets", "tracklet_offsets", "cand_offsets"):
        assert np.array_equal(f[k], g[k]), k
    assert ctx.flags() == 0
