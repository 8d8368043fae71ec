"""Serving loop: continuous batching under a memory budget; every served
event's triplets equal the oracle's on that event alone, and the HIP chain's
on that event alone."""
import concurrent.futures as cf

import numpy as np
import pytest

from helpers import rows_sorted

pytestmark = pytest.mark.gpu


def _oracle_triplets(sc, reqs, threads=16):
    """Per request: the oracle on its event alone, as input-hit indices."""
    from oracle import bridge

    def one(r):
        o = bridge.run_chain(sc, r.batch)
        return o["orig_index"][o["tracklets"]].astype(np.int64)
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, reqs))


def _keys(t):
    """Multiset canonical form of (n, 3) event-relative hit indices (< 2^21)."""
    t = np.asarray(t, np.uint64).reshape(-1, 3)
    return np.sort((t[:, 0] << np.uint64(42)) | (t[:, 1] << np.uint64(21)) | t[:, 2])


def _check_served(sc, reqs):
    want = _oracle_triplets(sc, reqs)
    for r, o in zip(reqs, want):
        assert r.triplets is not None, r.rid
        assert np.array_equal(_keys(r.triplets), _keys(o)), f"request {r.rid}: served triplets differ from the oracle"
    return sum(len(o) for o in want)


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
    assert _check_served(w.seed_config, reqs) > 0


@pytest.mark.timeout(600)
def test_serving_at_full_shape_matches_oracle():
    """The serving configuration at its stated shape: a 64 GiB device-memory
    budget, dynamic batches of up to 512 events, a Poisson stream of the
    70/30 pile-up-20 / pile-up-200 mix at a rate far above what the device
    serves (so batches fill up: more than 100 events each), and every one of
    the served events checked against the oracle on that event alone."""
    from trax_amd import events as ev
    from trax_amd import serving
    n = 1024
    w = ev.config_workload(4, n)
    reqs = serving.poisson_requests(w, n, rate_hz=5e6)
    srv = serving.SeedServer(w.seed_config, budget_bytes=64 << 30, max_batch=512, n_slots=3,
                             slot_hits=2_000_000)
    stats = srv.serve(reqs)
    print(f"served {stats.n_events} events in {stats.batches} batches (max {stats.max_batch}, mean "
          f"{stats.mean_batch:.0f}) on {stats.slots} slots, {stats.events_per_sec:.0f} events/s, "
          f"p50 {stats.p50_ms:.2f} ms p99 {stats.p99_ms:.2f} ms, {stats.used_bytes / 2**30:.1f} GiB", flush=True)
    assert stats.n_events == n and stats.slots == 3
    assert stats.used_bytes <= 64 << 30
    assert stats.max_batch > 100 and max(stats.batch_sizes) <= 512
    n_trk = _check_served(w.seed_config, reqs)
    print(f"{n} served events equal the oracle ({n_trk} triplets)", flush=True)


def test_slot_regrows_until_nothing_overflows():
    """A server whose slots start with tiny pair and triplet capacities: the
    first run overflows the pairs, the rerun (triplet capacity still sized
    from truncated pairs) the triplets, and the slot grows its buffers
    round after round until the batch fits; results match the oracle."""
    from trax_amd import events as ev
    from trax_amd import serving
    w = ev.config_workload(4, 12)
    reqs = serving.poisson_requests(w, 12, rate_hz=1e6)
    srv = serving.SeedServer(w.seed_config, budget_bytes=8 << 30, max_batch=12, n_slots=1,
                             slot_hits=200_000, pairs_per_hit=1, triplets_per_pair=1e-4)
    stats = srv.serve(reqs, realtime=False)
    assert stats.n_events == 12
    assert srv.slots[0].regrowths >= 2
    assert _check_served(w.seed_config, reqs) > 0


def test_budget_limits_slots():
    from trax_amd import events as ev
    from trax_amd import runtime, serving
    w = ev.config_workload(4, 4)
    with pytest.raises(runtime.TraxError):
        serving.SeedServer(w.seed_config, budget_bytes=1 << 20, max_batch=4, n_slots=2,
                           slot_hits=100_000)
