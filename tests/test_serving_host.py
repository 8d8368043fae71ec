"""Host side of the serving loop (no GPU): a request's packed blocks and the
batch assembly equal the chain's multi-event layout that
events.concat_events builds (the reference's HitCollection layout across the
events of a group, HitCollection.cpp:106-146)."""
import numpy as np


def test_packed_batch_equals_concatenated_events():
    from trax_amd import events as ev
    from trax_amd import serving as S
    w = ev.config_workload(4, 7)
    reqs = [S.Request(k, 0.0, ev.make_batch(w, events=[k])) for k in range(7)]
    reqs[3] = S.Request(3, 0.0, ev.EventBatch(*[np.zeros(0, np.float32)] * 3, *[np.zeros(0, np.uint32)] * 4,
                                              np.zeros(1, np.uint32), np.zeros(w.seed_config.n_layers, np.uint32),
                                              np.zeros(w.seed_config.n_layers, np.uint32),
                                              w.seed_config.n_layers))  # an empty event
    L = w.seed_config.n_layers
    n_max = sum(r.batch.n_hits for r in reqs) + 5
    hits = np.full((6, n_max), 0xDEADBEEF, np.uint32)
    evt = np.full(n_max, -1, np.int32)
    lay = np.zeros((2, 7 * L + 3), np.uint32)
    eoff = np.zeros(9, np.int32)
    n, offsets = S.pack_requests(reqs, hits, evt, lay, eoff)
    want = ev.concat_events([r.batch for r in reqs])
    assert n == want.n_hits
    for i, k in enumerate(S.HIT_ROWS):
        col = getattr(want, k)
        assert np.array_equal(hits[i, :n], col.view(np.uint32) if col.dtype == np.float32 else col), k
    assert np.array_equal(evt[:n], want.evt.astype(np.int32))
    assert np.array_equal(eoff[:7], want.event_offset.astype(np.int32))
    assert np.array_equal(offsets, want.event_offset.astype(np.int64))
    assert np.array_equal(lay[0, :7 * L], want.layer_nhits) and np.array_equal(lay[1, :7 * L], want.layer_offset)
    assert np.all(hits[:, n:] == 0xDEADBEEF)  # nothing written past the batch
