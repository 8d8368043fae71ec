"""GPU parity of the second pair generator, PairGeneratorSector
(PairGeneratorSector.h:43-234), through the C ABI: alone, in the whole chain
(trax_seed_chain with pair_generator = sector) and through the C++ class,
against the oracle's statement-by-statement restatement."""
import os
import subprocess

import numpy as np
import pytest

from helpers import cand_rows, rows_sorted, with_segment, written_pairs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "trax_amd", "_build", "trax_seed")
UNWRITTEN = 0xFFFFFFFF


def _workload(config_id, n_events, n_phi=None, spread_phi=None):
    from trax_amd import events as ev
    w = ev.config_workload(config_id, n_events)
    sc = w.seed_config
    sc.pair_generator = "sector"
    if n_phi is not None:
        sc.n_phi = n_phi
    if spread_phi is not None:
        for t in sc.triplets:
            t.pair_spread_phi = spread_phi
    return w, sc, ev.make_batch(w)


@pytest.mark.parametrize("config_id,n_events", [(0, 1), (1, 2), (2, 3)])
def test_sector_chain_matches_oracle(ctx, config_id, n_events):
    from oracle import bridge
    from trax_amd import runtime
    _, sc, batch = _workload(config_id, n_events)
    print(f"sector chain: config {config_id}, {batch.n_events} events, {batch.n_hits} hits", flush=True)
    bufs, _ = runtime.run_batch(ctx, sc, batch)
    torch.cuda.synchronize()
    g = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    o = bridge.run_chain(sc, batch)
    assert o["n_pairs_unwritten"] == 0
    assert np.array_equal(g["grid"], o["grid"])
    assert np.array_equal(g["pair_offsets"], o["pair_offsets"])
    gp = rows_sorted(with_segment(g["pairs"], g["pair_offsets"]))
    op_ = rows_sorted(with_segment(written_pairs(o["pairs"]), o["pair_offsets"]))
    assert len(gp) > 0 and np.array_equal(gp, op_)
    assert np.array_equal(g["cand_offsets"], o["cand_offsets"])
    assert np.array_equal(rows_sorted(cand_rows(g["pairs"], g["cands"], g["cand_offsets"])),
                          rows_sorted(cand_rows(o["pairs"], o["cands"], o["cand_offsets"])))
    assert np.array_equal(g["tracklet_offsets"], o["tracklet_offsets"])
    assert np.array_equal(rows_sorted(with_segment(g["tracklets"], g["tracklet_offsets"])),
                          rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"])))


def _pairs_only(ctx, sc, batch, two_phase):
    from trax_amd import runtime
    db = runtime.DeviceBatch.upload(batch)
    p, c, t = runtime.default_capacities(batch.n_hits)
    bufs = runtime.ChainBuffers(sc, batch.n_events, batch.n_hits, 64 * p, c, t)
    ctx.flags(clear=True)
    runtime.grid_build(ctx, bufs, db)
    if two_phase:
        runtime.pairs_count(ctx, bufs, batch.n_events)
        runtime.pairs_store(ctx, bufs, batch.n_events)
    else:
        runtime.pairs_sector(ctx, bufs, batch.n_events)
    torch.cuda.synchronize()
    assert ctx.flags() == 0
    n = runtime.totals(bufs)[0]
    K1 = batch.n_events * len(sc.triplets) + 1
    return (bufs.pairs[:2 * n].cpu().numpy().view(np.uint32).reshape(-1, 2),
            bufs.pair_offsets[:K1].cpu().numpy().view(np.uint32))


def test_sector_two_phase_equals_single_call(ctx):
    _, sc, batch = _workload(1, 2)
    a, ao = _pairs_only(ctx, sc, batch, two_phase=False)
    b, bo = _pairs_only(ctx, sc, batch, two_phase=True)
    assert np.array_equal(ao, bo) and np.array_equal(a, b)


def test_sector_duplicate_visits(ctx):
    """2 * spread + 1 = 9 phi cells on an 8-cell ring: the reference visits one
    column twice and keeps one bit; the pair sets agree, and each segment's
    GPU count is the reference's written count (its offsets also count the
    repeated visits, whose slots stay unwritten)."""
    from oracle import bridge
    _, sc, batch = _workload(1, 1, n_phi=8, spread_phi=4)
    pairs, offs = _pairs_only(ctx, sc, batch, two_phase=False)
    o = bridge.run_chain(sc, batch, stop_after=1)
    assert o["n_pairs_unwritten"] > 0
    assert np.array_equal(rows_sorted(with_segment(pairs, offs)),
                          rows_sorted(with_segment(written_pairs(o["pairs"]), o["pair_offsets"])))
    oo = o["pair_offsets"]
    written = np.array([(o["pairs"][oo[k]:oo[k + 1], 0] != UNWRITTEN).sum() for k in range(len(oo) - 1)])
    assert np.array_equal(np.diff(offs.astype(np.int64)), written)


def test_sector_rejects_window_wider_than_row(ctx):
    from trax_amd import runtime
    _, sc, batch = _workload(1, 1, n_phi=8, spread_phi=5)
    with pytest.raises(runtime.TraxError, match="pair_spread_phi"):
        runtime.run_batch(ctx, sc, batch)


def test_sector_cpp_class_matches_oracle(tmp_path):
    from oracle import bridge
    from trax_amd import eventfile
    _, sc, batch = _workload(1, 2)
    fin, fout = str(tmp_path / "events.bin"), str(tmp_path / "results.bin")
    eventfile.write_events(fin, sc, batch)
    print("trax_seed sector: config 1, 2 events", flush=True)
    r = subprocess.run([DRIVER, fin, fout, "sector"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = eventfile.read_results(fout, batch.n_events, len(sc.triplets))
    o = bridge.run_chain(sc, batch)
    assert np.array_equal(g["pair_offsets"], o["pair_offsets"])
    assert np.array_equal(rows_sorted(with_segment(g["pairs"], g["pair_offsets"])),
                          rows_sorted(with_segment(written_pairs(o["pairs"]), o["pair_offsets"])))
    assert np.array_equal(g["tracklet_offsets"], o["tracklet_offsets"])
    assert np.array_equal(rows_sorted(with_segment(g["tracklets"], g["tracklet_offsets"])),
                          rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"])))
