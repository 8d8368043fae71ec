"""Serving loop: continuous batching under a memory budget; every event's
triplets are the same served in the mix as computed alone."""
import numpy as np
import pytest

from helpers import rows_sorted

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(ctx):
    from trax_amd import events as ev
    from trax_amd import runtime, serving
    w = ev.config_workload(4, 24)  # 70/30 mix of pile-up 20 and 200
    reqs = serving.poisson_requests(w, 24, rate_hz=2000.0)
    srv = serving.SeedServer(w.seed_config, budget_bytes=8 << 30, max_batch=8, n_slots=3,
                             slot_hits=200_000)
    stats = srv.serve(reqs)
    print(f"served {stats.n_events} events in {stats.batches} batches on {stats.slots} slots, "
          f"p50 {stats.p50_ms:.2f} ms p99 {stats.p99_ms:.2f} ms", flush=True)
    assert stats.n_events == len(reqs) and stats.slots >= 2
    assert stats.used_bytes <= stats.budget_bytes
    for r in reqs:
        bufs, _ = runtime.run_batch(ctx, w.seed_config, r.batch)
        alone = runtime.fetch(bufs, 1, r.batch.n_hits)
        t = alone["orig_index"][alone["tracklets"]].astype(np.int64)
        assert np.array_equal(rows_sorted(r.triplets.astype(np.uint64)), rows_sorted(t.astype(np.uint64)))


def test_budget_limits_slots():
    from trax_amd import events as ev
    from trax_amd import runtime, serving
    w = ev.config_workload(4, 4)
    with pytest.raises(runtime.TraxError):
        serving.SeedServer(w.seed_config, budget_bytes=1 << 20, max_batch=4, n_slots=2,
                           slot_hits=100_000)
