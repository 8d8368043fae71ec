"""GPU parity of TripletConnectivity (TripletConnectivity.h:56-108,
.cpp:21-71) through the C ABI and the C++ class: the reference's known
answers, random sets against the oracle (small) and the numpy join (large),
hub keys, the all-ones tight key, modulo-2^32 weights, aliased sets,
workspace reuse, and chains over the triplets the seeding chain finds in the
real ttbar events."""
import os
import subprocess

import numpy as np
import pytest

from test_oracle_connectivity import KATS, as_set, chain, numpy_join, oracle_fn, random_sets

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "trax_amd", "_build", "trax_connect")
CMS = os.path.join(ROOT, "tests", "golden", "cms")


def _dev(a):
    from trax_amd import runtime
    return runtime._dev_u32(np.asarray(a, np.uint32), "cuda")


def _dev_set(s):
    from trax_amd import runtime
    return runtime.Tracklets(_dev(s[0]), _dev(s[1]), _dev(s[2]), _dev(s[3]))


def _host(t):
    return t.cpu().numpy().view(np.uint32).copy()


def gpu_fn(ctx):
    from trax_amd import runtime

    def fn(b, f, backward, tight):
        db, df = _dev_set(b), _dev_set(f)
        runtime.triplet_connectivity(ctx, db, df, backward=backward, tight=tight)
        torch.cuda.synchronize()
        return _host(df.connectivity)
    return fn


@pytest.mark.parametrize("name", sorted(KATS))
def test_known_answers_c_abi(ctx, name):
    from trax_amd import runtime
    sets, backward, tight, want = KATS[name]
    print(f"connectivity known answer {name}", flush=True)
    dev = [_dev_set(as_set(r)) for r in sets]
    for k in range(len(dev) - 1):  # the sets stay on the device between runs
        runtime.triplet_connectivity(ctx, dev[k], dev[k + 1], backward=backward, tight=tight)
    torch.cuda.synchronize()
    assert _host(dev[-1].connectivity).tolist() == want
    assert ctx.flags() == 0


@pytest.mark.parametrize("name", sorted(KATS))
def test_known_answers_cpp_class(tmp_path, name):
    sets, backward, tight, want = KATS[name]
    fin, fout = str(tmp_path / "sets.bin"), str(tmp_path / "con.bin")
    with open(fin, "wb") as f:
        np.array([0x43585254, 1, len(sets), int(backward) | (int(tight) << 1)], np.uint32).tofile(f)
        for rows in sets:
            s = as_set(rows)
            np.array([len(s[0])], np.uint32).tofile(f)
            for col in s:
                col.astype(np.uint32).tofile(f)
    print(f"trax_connect {name}", flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    con = np.fromfile(fout, np.uint32)
    assert con[-len(want):].tolist() == want
    assert np.array_equal(con, np.concatenate(chain(oracle_fn, sets, backward, tight)))


MODES = [(False, False), (True, False), (False, True)]


@pytest.mark.parametrize("backward,tight", MODES)
def test_random_sets_match_oracle(ctx, backward, tight):
    from oracle import bridge
    rng = np.random.default_rng(21)
    for nb, nf, hits in [(2000, 3000, 400), (1, 500, 3), (600, 1, 2), (64, 64, 64)]:
        b, f = random_sets(rng, nb, nf, hits)
        if tight:
            b[1] %= np.uint32(5)
            f[0] %= np.uint32(5)
        print(f"connectivity random: backward={backward} tight={tight} nb={nb} nf={nf}", flush=True)
        assert np.array_equal(gpu_fn(ctx)(b, f, backward, tight), bridge.connectivity(b, f, backward, tight))


@pytest.mark.parametrize("backward,tight", MODES)
def test_large_sets_match_numpy_join(ctx, backward, tight):
    rng = np.random.default_rng(22)
    b, f = random_sets(rng, 400_000, 700_000, 150_000, con_max=1000)
    if tight:  # about 400k distinct (hit2, hit3) keys, so most following triplets match
        b[1] %= np.uint32(97)
        f[0] %= np.uint32(97)
        b[2] %= np.uint32(4096)
        f[1] %= np.uint32(4096)
    print(f"connectivity large: backward={backward} tight={tight}", flush=True)
    got = gpu_fn(ctx)(b, f, backward, tight)
    want = numpy_join(b, f, backward, tight)
    assert np.array_equal(got, want)
    assert (got != f[3]).mean() > 0.5


def test_hub_key_all_ones_key_and_wrapping_weights(ctx):
    n = 60_000
    b = [np.zeros(n, np.uint32), np.full(n, 0xFFFFFFFF, np.uint32), np.full(n, 0xFFFFFFFF, np.uint32),
         np.arange(n, dtype=np.uint32)]
    b[1][: n // 2] = 7  # half the tight keys are (7, all-ones), half (all-ones, all-ones)
    b[3][-3:] = 0xFFFFFFFF  # con + 1 wraps to 0
    f = [np.array([0xFFFFFFFF, 7, 7, 3], np.uint32), np.array([0xFFFFFFFF, 0xFFFFFFFF, 3, 0], np.uint32),
         np.zeros(4, np.uint32), np.array([1, 2, 3, 4], np.uint32)]
    for tight in (False, True):
        got = gpu_fn(ctx)(b, f, False, tight)
        assert np.array_equal(got, numpy_join(b, f, False, tight)), tight
    # wide: every base hit3 is all-ones: the first following triplet collects them all
    assert gpu_fn(ctx)(b, f, False, False)[0] == (1 + sum(int(c) + 1 for c in b[3])) % 2**32


def test_aliased_base_and_following(ctx):
    from trax_amd import runtime
    rng = np.random.default_rng(23)
    s, _ = random_sets(rng, 200_000, 1, 50_000)
    want = numpy_join([c.copy() for c in s], s, False, False)
    d = _dev_set(s)
    runtime.triplet_connectivity(ctx, d, d)
    torch.cuda.synchronize()
    assert np.array_equal(_host(d.connectivity), want)


def test_workspace_reuse_with_smaller_second_call(ctx):
    from trax_amd import runtime
    rng = np.random.default_rng(24)
    ws = runtime.connectivity_workspace(ctx, 250_000)
    for nb, nf in [(250_000, 300_000), (1_000, 2_000)]:
        b, f = random_sets(rng, nb, nf, 20_000)
        db, df = _dev_set(b), _dev_set(f)
        runtime.triplet_connectivity(ctx, db, df, tight=False, workspace=ws)
        torch.cuda.synchronize()
        assert np.array_equal(_host(df.connectivity), numpy_join(b, f, False, False))


def test_rejected_arguments(ctx):
    from trax_amd import runtime
    s = _dev_set(as_set([(1, 2, 3)]))
    with pytest.raises(runtime.TraxError, match="forward only"):
        runtime.triplet_connectivity(ctx, s, s, backward=True, tight=True)
    small = torch.empty(4, dtype=torch.int32, device="cuda")
    with pytest.raises(runtime.TraxError, match="workspace"):
        runtime.triplet_connectivity(ctx, s, s, workspace=small)
    missing = runtime.Tracklets(None, None, None, s.connectivity)
    with pytest.raises(runtime.TraxError, match="column"):
        runtime.triplet_connectivity(ctx, missing, s)


def test_ttbar_triplet_chains(ctx):
    """The seeding chain's triplets on the 10 ttbar events, per layer
    triplet: tight chains 0 -> 1 -> 2 -> 3 (triplets sharing two hits) and
    wide chains 0 -> 2 (triplet 0's outer hit is triplet 2's inner hit), on
    the GPU against the oracle; some chains exist."""
    from oracle import bridge
    from trax_amd import cmsdata, pbio, runtime
    evs = pbio.read_event_container(os.path.join(CMS, "hits.ttBar.pb"))
    g = cmsdata.Geometry.load(os.path.join(CMS, "detectorRadius.dat"), os.path.join(CMS, "radiusDictionary.dat"))
    batch, _ = cmsdata.load_events(evs, g, cmsdata.reference_triplets(), 1.0)
    sc = cmsdata.ttbar_workload(g)
    bufs, _ = runtime.run_batch(ctx, sc, batch)
    torch.cuda.synchronize()
    out = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    t, off, P = out["tracklets"], out["tracklet_offsets"], len(sc.triplets)
    per = []
    for p in range(P):
        idx = np.concatenate([np.arange(off[e * P + p], off[e * P + p + 1]) for e in range(batch.n_events)])
        rows = t[idx.astype(np.int64)]
        per.append([rows[:, 0].copy(), rows[:, 1].copy(), rows[:, 2].copy(), np.zeros(len(rows), np.uint32)])
    dev = [_dev_set(s) for s in per]
    for k in range(P - 1):
        runtime.triplet_connectivity(ctx, dev[k], dev[k + 1], tight=True)
    torch.cuda.synchronize()
    host = [[c.copy() for c in s] for s in per]
    for k in range(P - 1):
        host[k + 1][3] = bridge.connectivity(host[k], host[k + 1], False, True)
    for k in range(P):
        assert np.array_equal(_host(dev[k].connectivity), host[k][3]), k
    assert host[P - 1][3].max() >= 3  # some four-layer chains 1-2-3-4-5-8
    wide = _dev_set(per[2])
    runtime.triplet_connectivity(ctx, dev[0], wide)
    torch.cuda.synchronize()
    w = bridge.connectivity(host[0], per[2], False, False)
    assert np.array_equal(_host(wide.connectivity), w) and w.sum() > 0
