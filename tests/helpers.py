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
