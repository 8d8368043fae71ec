"""The C++ algorithm classes (GridBuilder, PairGeneratorBeamspot,
TripletThetaPhiPredictor, TripletThetaPhiFilter) driven by trax_seed, in the
reference's stage order, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from helpers import cand_rows, rows_sorted, with_segment, written_pairs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "trax_amd", "_build", "trax_seed")


@pytest.mark.parametrize("config_id,n_events", [(0, 1), (3, 2)])
def test_cpp_classes_match_oracle(tmp_path, config_id, n_events):
    from oracle import bridge
    from trax_amd import eventfile
    from trax_amd import events as ev
    w = ev.config_workload(config_id, n_events)
    sc, batch = w.seed_config, ev.make_batch(w)
    fin, fout = str(tmp_path / "events.bin"), str(tmp_path / "results.bin")
    eventfile.write_events(fin, sc, batch)
    print(f"trax_seed config {config_id}, {n_events} events", flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = eventfile.read_results(fout, batch.n_events, len(sc.triplets))
    o = bridge.run_chain(sc, batch)
    assert np.array_equal(g["orig_index"], o["orig_index"])
    assert np.array_equal(g["grid"], o["grid"])
    assert np.array_equal(g["pair_offsets"], o["pair_offsets"])
    assert np.array_equal(rows_sorted(with_segment(g["pairs"], g["pair_offsets"])),
                          rows_sorted(with_segment(written_pairs(o["pairs"]), o["pair_offsets"])))
    assert np.array_equal(g["cand_offsets"], o["cand_offsets"])
    assert np.array_equal(rows_sorted(cand_rows(g["pairs"], g["cands"], g["cand_offsets"])),
                          rows_sorted(cand_rows(o["pairs"], o["cands"], o["cand_offsets"])))
    assert np.array_equal(g["tracklet_offsets"], o["tracklet_offsets"])
    assert np.array_equal(rows_sorted(with_segment(g["tracklets"], g["tracklet_offsets"])),
                          rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"])))


@pytest.mark.parametrize("config_id,n_events", [(1, 2), (3, 2)])
def test_cpp_fused_pass_equals_staged_classes(tmp_path, config_id, n_events):
    """TripletThetaPhiPredictFilter (trax_seed ... fused) writes the same
    triplets, in the same order, and the same triplet and candidate offsets
    as TripletThetaPhiPredictor + TripletThetaPhiFilter, and no candidate
    records."""
    from trax_amd import eventfile
    from trax_amd import events as ev
    w = ev.config_workload(config_id, n_events)
    sc, batch = w.seed_config, ev.make_batch(w)
    fin = str(tmp_path / "events.bin")
    eventfile.write_events(fin, sc, batch)
    res = {}
    for mode in ("staged", "fused"):
        fout = str(tmp_path / f"{mode}.bin")
        print(f"trax_seed {mode}: config {config_id}, {n_events} events", flush=True)
        args = [DRIVER, fin, fout] + (["beamspot", "fused"] if mode == "fused" else [])
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        res[mode] = eventfile.read_results(fout, batch.n_events, len(sc.triplets))
    s, f = res["staged"], res["fused"]
    assert len(f["cands"]) == 0 and len(s["cands"]) > 0
    for k in ("grid", "pairs", "pair_offsets", "cand_offsets", "tracklets", "tracklet_offsets"):
        assert np.array_equal(s[k], f[k]), k
    assert len(f["tracklets"]) > 0


@pytest.mark.parametrize("config_id,n_events,gen", [(3, 2, "beamspot"), (1, 2, "sector")])
def test_cpp_graph_replays_match_oracle(tmp_path, config_id, n_events, gen):
    """SeedChainGraph (trax_seed ... graph): the whole chain recorded once as
    a HIP graph through trax_chain_graph_capture and replayed three times by
    a C++ caller with no PyTorch; the driver checks every replay against the
    stage-by-stage classes, and the last replay's results match the oracle."""
    from oracle import bridge
    from trax_amd import eventfile
    from trax_amd import events as ev
    w = ev.config_workload(config_id, n_events)
    sc, batch = w.seed_config, ev.make_batch(w)
    sc.pair_generator = gen
    fin, fout = str(tmp_path / "events.bin"), str(tmp_path / "results.bin")
    eventfile.write_events(fin, sc, batch)
    print(f"trax_seed {gen} graph: config {config_id}, {n_events} events", flush=True)
    r = subprocess.run([DRIVER, fin, fout, gen, "graph"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = eventfile.read_results(fout, batch.n_events, len(sc.triplets))
    o = bridge.run_chain(sc, batch)
    assert len(g["cands"]) == 0
    assert np.array_equal(g["grid"], o["grid"])
    assert np.array_equal(g["pair_offsets"], o["pair_offsets"])
    assert np.array_equal(rows_sorted(with_segment(g["pairs"], g["pair_offsets"])),
                          rows_sorted(with_segment(written_pairs(o["pairs"]), o["pair_offsets"])))
    assert np.array_equal(g["cand_offsets"], o["cand_offsets"])
    assert np.array_equal(g["tracklet_offsets"], o["tracklet_offsets"])
    assert np.array_equal(rows_sorted(with_segment(g["tracklets"], g["tracklet_offsets"])),
                          rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"])))
    assert len(o["tracklets"]) > 0
