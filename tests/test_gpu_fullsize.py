"""Full-size batches against the oracle: the 64-event pile-up-200 batch the
benchmark times and the 256-event uneven pile-up-20..200 batch, every event
checked (sorted columns, grid, segment sizes, pair / candidate / triplet
multisets per segment) against the oracle run on that event alone, with the
batch offsets equal to the running sums of those segment sizes.  The fused
prediction + filter pass must give the staged chain's arrays."""
import numpy as np
import pytest

from helpers import compare_batch_per_event, oracle_per_event

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
@pytest.mark.parametrize("config_id", [3, 2])
def test_full_batch_matches_oracle_event_by_event(ctx, config_id):
    import torch

    from trax_amd import events as ev
    from trax_amd import runtime
    w = ev.config_workload(config_id)
    sc, batch = w.seed_config, ev.make_batch(w)
    print(f"full batch {w.name}: {batch.n_events} events, {batch.n_hits} hits, grid {sc.n_phi}x{sc.n_z}",
          flush=True)
    bufs, _ = runtime.run_batch(ctx, sc, batch)
    torch.cuda.synchronize()
    g = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    del bufs
    fb, _ = runtime.run_batch(ctx, sc, batch, candidates=False)
    torch.cuda.synchronize()
    f = runtime.fetch(fb, batch.n_events, batch.n_hits)
    del fb
    for k in ("grid", "orig_index", "pairs", "pair_offsets", "cand_offsets", "tracklets", "tracklet_offsets"):
        assert np.array_equal(f[k], g[k]), k
    print(f"GPU: {len(g['pairs'])} pairs, {len(g['cands'])} candidate slots, {len(g['tracklets'])} triplets",
          flush=True)
    o = oracle_per_event(sc, batch)
    print("oracle done", flush=True)
    compare_batch_per_event(g, batch, len(sc.triplets), o)
    assert len(g["tracklets"]) > 100_000
