"""Launch shapes change nothing but speed; the autotuner's pick stays bit-exact."""
import numpy as np
import pytest

from helpers import rows_sorted, with_segment

pytestmark = pytest.mark.gpu


def _trk(ctx, sc, batch):
    from trax_amd import runtime
    bufs, _ = runtime.run_batch(ctx, sc, batch)
    g = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    return g


def _plain(ctx, sc, batch):
    """Two-phase stage calls (the plain method's kernels) on the same batch."""
    from trax_amd import runtime
    bufs, db = runtime.run_batch(ctx, sc, batch)
    E = batch.n_events
    runtime.grid_build(ctx, bufs, db)
    runtime.pairs_count(ctx, bufs, E)
    runtime.pairs_store(ctx, bufs, E)
    runtime.predict_count(ctx, bufs, E)
    runtime.predict_store(ctx, bufs, E)
    runtime.triplet_filter(ctx, bufs, E)
    assert ctx.flags() == 0
    return runtime.fetch(bufs, batch.n_events, batch.n_hits)


def test_launch_shapes_do_not_change_results(ctx):
    from trax_amd import events as ev
    w = ev.config_workload(2, 3)
    batch = ev.make_batch(w)
    ref = None
    try:
        for waves, pb, qb, tile in [(4, 256, 256, 8192), (1, 64, 512, 4096), (8, 512, 64, 16384),
                                    (2, 128, 128, 8192)]:
            print(f"launch {waves} {pb} {qb} {tile}", flush=True)
            ctx.set_launch(waves, pb, qb, tile)
            g = _trk(ctx, w.seed_config, batch)
            plain = _plain(ctx, w.seed_config, batch)
            key = {k: g[k] for k in ("grid", "orig_index", "pair_offsets", "cand_offsets",
                                     "tracklet_offsets")}
            key["pairs"] = rows_sorted(g["pairs"])
            key["trk"] = rows_sorted(with_segment(g["tracklets"], g["tracklet_offsets"]))
            for k in ("pair_offsets", "cand_offsets", "tracklet_offsets"):
                assert np.array_equal(plain[k], g[k]), k
            assert np.array_equal(rows_sorted(plain["pairs"]), key["pairs"])
            assert np.array_equal(rows_sorted(with_segment(plain["tracklets"], plain["tracklet_offsets"])),
                                  key["trk"])
            if ref is None:
                ref = key
            else:
                for k in ref:
                    assert np.array_equal(ref[k], key[k]), k
    finally:
        ctx.set_launch()  # library defaults


def test_autotuned_binning_matches_oracle(ctx):
    from oracle import bridge
    from trax_amd import autotune
    from trax_amd import events as ev
    w = ev.config_workload(2, 2)
    batch = ev.make_batch(w)
    try:
        res = autotune.tune(ctx, w, batch, phi_bins=(64, 128), z_bins=(32, 64), reps=1)
        assert (res.n_phi, res.n_z) in {(64, 32), (64, 64), (128, 32), (128, 64)}
        autotune.apply(ctx, w, res)
        g = _trk(ctx, w.seed_config, batch)
    finally:
        ctx.set_launch()  # library defaults
    o = bridge.run_chain(w.seed_config, batch)
    assert np.array_equal(g["grid"], o["grid"])
    assert np.array_equal(g["tracklet_offsets"], o["tracklet_offsets"])
    assert np.array_equal(rows_sorted(with_segment(g["tracklets"], g["tracklet_offsets"])),
                          rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"])))


def test_fused_walk_shapes_do_not_change_results(ctx):
    """Every compiled shape of the fused prediction + filter walk (lane buffer
    depth, waves per block, waves per SIMD) writes the same triplets,
    offsets and totals, on pile-up events and on wide windows that fill the
    spill pools."""
    from trax_amd import events as ev
    from trax_amd import runtime
    cases = [ev.config_workload(3, 2), ev.config_workload(1, 1)]
    for t in cases[1].seed_config.triplets:
        t.sigma_phi, t.sigma_z = 1.5, 0.1
    keys = ("pairs", "pair_offsets", "cand_offsets", "tracklets", "tracklet_offsets")
    try:
        for w in cases:
            batch = ev.make_batch(w)
            ref = None
            for shape in runtime.FUSED_SHAPES:
                print(f"{w.name}: fused shape {shape}", flush=True)
                ctx.set_launch(fused_buffer=shape[0], fused_block_waves=shape[1], fused_waves_per_eu=shape[2])
                bufs, _ = runtime.run_batch(ctx, w.seed_config, batch, candidates=False)
                g = runtime.fetch(bufs, batch.n_events, batch.n_hits)
                if ref is None:
                    ref = g
                    assert len(g["tracklets"]) > 0
                for k in keys:
                    assert np.array_equal(g[k], ref[k]), (shape, k)
        with pytest.raises(runtime.TraxError):
            ctx.set_launch(fused_buffer=5, fused_block_waves=4, fused_waves_per_eu=7)
    finally:
        ctx.set_launch()
