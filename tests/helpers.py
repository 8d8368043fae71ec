"""Comparison helpers shared by the parity tests (test code, not product)."""
import numpy as np

UNWRITTEN = 0xFFFFFFFF


def rows_sorted(a: np.ndarray) -> np.ndarray:
    """Rows of a 2-D uint32 array in lexicographic order (multiset canonical form)."""
    a = np.asarray(a, dtype=np.uint64)
    if a.size == 0:
        return a.reshape(0, a.shape[1] if a.ndim == 2 else 0)
    keys = tuple(a[:, i] for i in reversed(range(a.shape[1])))
    return a[np.lexsort(keys)]


def written_pairs(pairs: np.ndarray) -> np.ndarray:
    return pairs[pairs[:, 0] != UNWRITTEN]


def cand_triples(pairs: np.ndarray, cands: np.ndarray) -> np.ndarray:
    """(hit1, hit2, hit3) of every written candidate."""
    c = cands[cands[:, 0] != UNWRITTEN]
    p = pairs[c[:, 0]]
    return np.stack([p[:, 0], p[:, 1], c[:, 1]], axis=1)


def segment_ids(offsets: np.ndarray, n: int) -> np.ndarray:
    """Segment index of every element 0..n-1 given K+1 offsets."""
    return np.searchsorted(offsets, np.arange(n), side="right") - 1


def with_segment(rows: np.ndarray, offsets: np.ndarray) -> np.ndarray:
    seg = segment_ids(offsets, rows.shape[0]).astype(np.uint64)
    return np.concatenate([seg[:, None], rows.astype(np.uint64)], axis=1)


def to_input(rows: np.ndarray, orig_index: np.ndarray) -> np.ndarray:
    return orig_index[rows]


def cand_rows(pairs: np.ndarray, cands: np.ndarray, offsets: np.ndarray) -> np.ndarray:
    """(segment, hit1, hit2, hit3) of every written candidate."""
    seg = segment_ids(offsets, cands.shape[0]).astype(np.uint64)
    mask = cands[:, 0] != UNWRITTEN
    return np.concatenate([seg[mask][:, None], cand_triples(pairs, cands).astype(np.uint64)], axis=1)


# ------------------------------------------------- per-event batch comparison
def _event_keys(rows_in: np.ndarray, lt: np.ndarray) -> np.ndarray:
    """One uint64 per row of event-relative input-hit indices (< 2^16 each)
    and its layer triplet: a sort of these keys is the segment multiset."""
    rows_in = rows_in.astype(np.uint64)
    k = lt.astype(np.uint64) << np.uint64(48)
    for c in range(rows_in.shape[1]):
        k |= rows_in[:, c] << np.uint64(16 * (rows_in.shape[1] - 1 - c))
    return np.sort(k)


def _segment_rows(rows: np.ndarray, offsets: np.ndarray, e: int, P: int):
    """Rows of event e's segments and their layer triplet."""
    lo, hi = int(offsets[e * P]), int(offsets[(e + 1) * P])
    lt = np.repeat(np.arange(P), np.diff(offsets[e * P:(e + 1) * P + 1].astype(np.int64)))
    return rows[lo:hi], lt


def compare_batch_per_event(g: dict, batch, P: int, oracle_events, stages=("pairs", "cands", "tracklets")):
    """A batch result `g` (fetch() of the HIP chain, or the oracle on the
    whole batch) against the oracle run on every event alone: per event, the
    sorted columns, the grid (shifted by the event's first hit), the segment
    sizes of every stage, and the pair, candidate and triplet multisets of
    every (event, layer triplet) segment as event-relative input-hit indices.
    Batch offsets must be the running sums of the per-event segment sizes."""
    E = batch.n_events
    assert E == len(oracle_events)
    assert int(np.max(batch.layer_nhits.reshape(E, -1).sum(axis=1), initial=0)) < (1 << 16)
    grid_per_event = g["grid"].shape[0] // E
    for stage, off_key in (("pairs", "pair_offsets"), ("cands", "cand_offsets"), ("tracklets", "tracklet_offsets")):
        if stage not in stages:
            continue
        sizes = np.concatenate([np.diff(o[off_key].astype(np.int64)) for o in oracle_events])
        want = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
        assert np.array_equal(g[off_key], want), off_key
    for e, o in enumerate(oracle_events):
        base = int(batch.event_offset[e])
        n = o["x"].shape[0]
        sl = slice(base, base + n)
        for k in ("x", "y", "z"):
            assert np.array_equal(g[k][sl].view(np.uint32), o[k].view(np.uint32)), (e, k)
        for k in ("layer", "det_id", "hit_id"):
            assert np.array_equal(g[k][sl], o[k]), (e, k)
        assert np.all(g["evt"][sl] == e)
        assert np.array_equal(g["orig_index"][sl] - base, o["orig_index"]), e
        ge = g["grid"][e * grid_per_event:(e + 1) * grid_per_event]
        assert np.array_equal(ge - np.uint32(base), o["grid"]), e
        gin = g["orig_index"].astype(np.int64) - base  # sorted -> event-relative input index
        oin = o["orig_index"].astype(np.int64)
        if "pairs" in stages:
            gr, glt = _segment_rows(g["pairs"], g["pair_offsets"], e, P)
            orr, olt = _segment_rows(o["pairs"], o["pair_offsets"], 0, P)
            m = orr[:, 0] != UNWRITTEN
            assert np.array_equal(_event_keys(gin[gr], glt), _event_keys(oin[orr[m]], olt[m])), (e, "pairs")
        if "cands" in stages and g.get("cands") is not None:
            gr, glt = _segment_rows(g["cands"], g["cand_offsets"], e, P)
            orr, olt = _segment_rows(o["cands"], o["cand_offsets"], 0, P)
            gm, om = gr[:, 0] != UNWRITTEN, orr[:, 0] != UNWRITTEN
            assert gm.sum() == om.sum(), (e, "unwritten candidate slots")
            gt = np.stack([g["pairs"][gr[gm, 0], 0], g["pairs"][gr[gm, 0], 1], gr[gm, 1]], axis=1)
            ot = np.stack([o["pairs"][orr[om, 0], 0], o["pairs"][orr[om, 0], 1], orr[om, 1]], axis=1)
            assert np.array_equal(_event_keys(gin[gt], glt[gm]), _event_keys(oin[ot], olt[om])), (e, "cands")
        if "tracklets" in stages:
            gr, glt = _segment_rows(g["tracklets"], g["tracklet_offsets"], e, P)
            orr, olt = _segment_rows(o["tracklets"], o["tracklet_offsets"], 0, P)
            assert np.array_equal(_event_keys(gin[gr], glt), _event_keys(oin[orr], olt)), (e, "tracklets")


def oracle_per_event(sc, batch, threads: int = 16):
    """The oracle on every event of the batch alone, on a thread pool (the
    ctypes call releases the GIL)."""
    import concurrent.futures as cf

    from oracle import bridge
    from trax_amd import events as ev
    singles = [ev.concat_events([batch], subset=[e]) for e in range(batch.n_events)]
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda b: bridge.run_chain(sc, b), singles))
