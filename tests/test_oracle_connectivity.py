"""The oracle's TripletConnectivity restatement
(TripletConnectivity.h:56-108, .cpp:21-71) pinned by:

* the reference's own known answers (TripletConnectivityTest.h: test_run,
  test_run_backward, test_run_multilayer, test_run_tight_multilayer), chained
  the way those tests chain run() calls;
* a numpy join on random sets (group the base weights by key with
  np.unique + bincount, look the following keys up with searchsorted), which
  shares nothing with the oracle's all-pairs loop;
* the join order of the wide kernel as printed in the reference (base column
  indexed by the following triplet's id): it fails test_run_multilayer, which
  is why the oracle follows the tests.
"""
import numpy as np
import pytest

# (hit1, hit2, hit3) rows of TripletConnectivityTest.h (the id column dropped)
SIMPLE_FORWARD = [[(1, 2, 3), (1, 2, 4)], [(3, 5, 6), (9, 4, 8)]]
SIMPLE_BACKWARD = [[(3, 2, 23), (5, 2, 1)], [(12, 8, 3), (15, 11, 8)]]
FOUR_LAYER = [
    [(1, 3, 8), (1, 3, 5), (2, 4, 6), (2, 4, 7)],
    [(6, 11, 16), (9, 12, 17), (9, 13, 18), (9, 14, 19)],
    [(16, 21, 30), (15, 22, 29), (15, 24, 29), (15, 25, 27)],
    [(30, 31, 36), (28, 32, 37), (26, 33, 38), (27, 34, 39)],
]
FOUR_LAYER_TIGHT = [
    [(1, 4, 6), (1, 3, 5), (2, 3, 8), (2, 4, 7)],
    [(4, 6, 16), (9, 12, 17), (9, 13, 18), (9, 14, 19)],
    [(6, 16, 30), (15, 22, 29), (15, 24, 29), (15, 25, 27)],
    [(16, 30, 36), (28, 32, 37), (26, 33, 38), (27, 34, 39)],
]
# (sets, backward, tight, the asserted connectivity of the last set)
KATS = {
    "test_run": (SIMPLE_FORWARD, False, False, [1, 0]),
    "test_run_backward": (SIMPLE_BACKWARD, True, False, [1, 0]),
    "test_run_multilayer": (FOUR_LAYER, False, False, [3, 0, 0, 1]),
    "test_run_tight_multilayer": (FOUR_LAYER_TIGHT, False, True, [3, 0, 0, 0]),
}


def as_set(rows, con=None):
    a = np.array(rows, np.uint32).reshape(-1, 3)
    c = np.zeros(len(a), np.uint32) if con is None else np.asarray(con, np.uint32)
    return [a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy(), c]


def chain(fn, sets, backward, tight):
    """run(A, B), run(B, C), ... with fn(base, following, backward, tight) ->
    the following set's new connectivity."""
    s = [as_set(r) for r in sets]
    for k in range(len(s) - 1):
        s[k + 1][3] = fn(s[k], s[k + 1], backward, tight)
    return [x[3] for x in s]


def oracle_fn(b, f, backward, tight):
    from oracle import bridge
    return bridge.connectivity(b, f, backward, tight)


@pytest.mark.parametrize("name", sorted(KATS))
def test_reference_known_answers(name):
    sets, backward, tight, want = KATS[name]
    got = chain(oracle_fn, sets, backward, tight)
    assert got[-1].tolist() == want


def _printed_wide_kernel(b, f, backward, tight):
    """tripletConnectivityWide exactly as printed (TripletConnectivity.h:64-76):
    connected = baseHit[gid] == followHit[i], con[gid] += connected * (baseCon[i] + 1)."""
    assert not tight and len(b[0]) == len(f[0])
    bh = b[0] if backward else b[2]
    fh = f[2] if backward else f[0]
    out = f[3].copy()
    for gid in range(len(out)):
        for i in range(len(bh)):
            out[gid] += np.uint32(bh[gid] == fh[i]) * (b[3][i] + np.uint32(1))
    return out


def test_printed_wide_kernel_index_order_misses_the_known_answer():
    got = chain(_printed_wide_kernel, FOUR_LAYER, False, False)
    assert got[-1].tolist() == [2, 0, 0, 1]  # the test asserts [3, 0, 0, 1]
    # on the single-step tests both index orders agree
    assert chain(_printed_wide_kernel, SIMPLE_FORWARD, False, False)[-1].tolist() == [1, 0]
    assert chain(_printed_wide_kernel, SIMPLE_BACKWARD, True, False)[-1].tolist() == [1, 0]


def numpy_join(b, f, backward, tight):
    """Sum of (con + 1) per base key; each following key looks its sum up."""
    if tight:
        bk = (b[1].astype(np.uint64) << np.uint64(32)) | b[2].astype(np.uint64)
        fk = (f[0].astype(np.uint64) << np.uint64(32)) | f[1].astype(np.uint64)
    elif backward:
        bk, fk = b[0].astype(np.uint64), f[2].astype(np.uint64)
    else:
        bk, fk = b[2].astype(np.uint64), f[0].astype(np.uint64)
    keys, inv = np.unique(bk, return_inverse=True)
    w = (b[3].astype(np.uint64) + 1) & 0xFFFFFFFF
    sums = np.zeros(len(keys), np.uint64)
    np.add.at(sums, inv, w)
    pos = np.clip(np.searchsorted(keys, fk), 0, max(len(keys) - 1, 0))
    hit = (keys[pos] == fk) if len(keys) else np.zeros(len(fk), bool)
    add = np.where(hit, sums[pos] if len(keys) else 0, 0)
    return ((f[3].astype(np.uint64) + add) & 0xFFFFFFFF).astype(np.uint32)


def random_sets(rng, nb, nf, hits, con_max=5):
    b = [rng.integers(0, hits, nb, dtype=np.uint32) for _ in range(3)] + [
        rng.integers(0, con_max, nb, dtype=np.uint32)]
    f = [rng.integers(0, hits, nf, dtype=np.uint32) for _ in range(3)] + [
        rng.integers(0, con_max, nf, dtype=np.uint32)]
    return b, f


@pytest.mark.parametrize("backward,tight", [(False, False), (True, False), (False, True)])
def test_oracle_equals_numpy_join(backward, tight):
    rng = np.random.default_rng(11)
    for nb, nf, hits in [(1500, 2200, 300), (700, 400, 12), (50, 3000, 8)]:
        b, f = random_sets(rng, nb, nf, hits)
        if tight:  # tight keys need repeated (hit2, hit3) pairs to be interesting
            b[1] %= np.uint32(6)
            f[0] %= np.uint32(6)
        got = oracle_fn(b, f, backward, tight)
        want = numpy_join(b, f, backward, tight)
        assert np.array_equal(got, want)
        assert (got != f[3]).sum() > 0


def test_weights_wrap_modulo_2_32():
    """con = 0xFFFFFFFF adds con + 1 = 0; sums wrap."""
    b = as_set([(0, 0, 5), (0, 0, 5), (0, 0, 5)], [0xFFFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF])
    f = as_set([(5, 0, 0)], [7])
    assert oracle_fn(b, f, False, False).tolist() == [(7 + 2 * 0x80000000) % 2**32]


def test_aliased_sets_read_base_before_update():
    from oracle import bridge
    s = as_set([(1, 2, 3), (3, 4, 5), (5, 6, 7)], [0, 0, 0])
    got = bridge.connectivity(s, s, False, False)
    assert got.tolist() == [0, 1, 1]  # every base weight is 0 + 1, read before any update


def test_tight_backward_rejected_and_empty_sets():
    from oracle import bridge
    s = as_set([(1, 2, 3)])
    with pytest.raises(RuntimeError):
        bridge.connectivity(s, s, True, True)
    e = as_set(np.zeros((0, 3)))
    assert bridge.connectivity(e, s, False, False).tolist() == [0]
    assert bridge.connectivity(s, e, False, False).tolist() == []
