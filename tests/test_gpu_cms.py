"""GPU parity on the reference's real data: the 10 CMS ttbar events of
hits.ttBar.pb with the CMS geometry and the reference's ttbar configuration
(grid 50 x 8, minPt 1 GeV, the four layer triplets of layerTriplets.xml),
through the C ABI with both pair generators, against the oracle; and the
physics table computed from the GPU's triplets equals the oracle's."""
import os

import numpy as np
import pytest

from helpers import cand_rows, rows_sorted, with_segment, written_pairs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CMS = os.path.join(ROOT, "tests", "golden", "cms")


@pytest.fixture(scope="module")
def ttbar():
    from trax_amd import cmsdata, pbio
    evs = pbio.read_event_container(os.path.join(CMS, "hits.ttBar.pb"))
    g = cmsdata.Geometry.load(os.path.join(CMS, "detectorRadius.dat"), os.path.join(CMS, "radiusDictionary.dat"))
    batch, tracks = cmsdata.load_events(evs, g, cmsdata.reference_triplets(), 1.0)
    return evs, g, batch, tracks


@pytest.mark.parametrize("generator", ["beamspot", "sector"])
def test_ttbar_chain_matches_oracle(ctx, ttbar, generator):
    from oracle import bridge
    from trax_amd import cmsdata, physics, runtime
    evs, g, batch, tracks = ttbar
    sc = cmsdata.ttbar_workload(g)
    sc.pair_generator = generator
    print(f"ttbar chain ({generator}): {batch.n_events} events, {batch.n_hits} hits", flush=True)
    bufs, _ = runtime.run_batch(ctx, sc, batch)
    torch.cuda.synchronize()
    gpu = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    o = bridge.run_chain(sc, batch)
    assert o["n_pairs_unwritten"] == 0
    for k in ("x", "y", "z", "layer", "det_id", "hit_id", "evt", "orig_index", "grid"):
        assert np.array_equal(gpu[k], o[k]), k
    assert np.array_equal(gpu["pair_offsets"], o["pair_offsets"])
    assert np.array_equal(rows_sorted(with_segment(gpu["pairs"], gpu["pair_offsets"])),
                          rows_sorted(with_segment(written_pairs(o["pairs"]), o["pair_offsets"])))
    assert np.array_equal(gpu["cand_offsets"], o["cand_offsets"])
    assert np.array_equal(rows_sorted(cand_rows(gpu["pairs"], gpu["cands"], gpu["cand_offsets"])),
                          rows_sorted(cand_rows(o["pairs"], o["cands"], o["cand_offsets"])))
    assert np.array_equal(gpu["tracklet_offsets"], o["tracklet_offsets"])
    assert len(gpu["tracklets"]) > 0
    assert np.array_equal(rows_sorted(with_segment(gpu["tracklets"], gpu["tracklet_offsets"])),
                          rows_sorted(with_segment(o["tracklets"], o["tracklet_offsets"])))
    P = len(sc.triplets)
    assert physics.evaluate(gpu, evs, tracks, P).csv() == physics.evaluate(o, evs, tracks, P).csv()
