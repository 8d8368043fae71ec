"""Event I/O on the reference's real data: the PEventContainer / EventStore
codec (Event.proto:1-37, PEventStore.h:25-245), the loaders
(EventLoader.h:15-168), the CMS geometry (TripletBuilder.cpp:100-141,
DetectorGeometry.cpp:4-30) and HitCollection::addEvent's track selection and
layer ordering (HitCollection.cpp:20-146).

What pins them without the reference running:
* protobuf's published wire-format examples (150 -> 08 96 01; a negative
  enum is a ten-byte varint);
* the decoder re-encodes tests/golden/cms/hits.ttBar.pb byte for byte;
* every hit's detectorId resolves to a module of the geometry whose layer is
  the hit's own layer field, and hit radii lie in the module-radius band of
  their layer (two files written independently of the event file);
* the loader's selection and ordering against a numpy restatement built on
  lexsort and set coverage instead of the per-track loops.
"""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CMS = os.path.join(ROOT, "tests", "golden", "cms")
PB = os.path.join(CMS, "hits.ttBar.pb")


@pytest.fixture(scope="module")
def events():
    from trax_amd import pbio
    return pbio.read_event_container(PB)


@pytest.fixture(scope="module")
def geom():
    from trax_amd import cmsdata
    return cmsdata.Geometry.load(os.path.join(CMS, "detectorRadius.dat"),
                                 os.path.join(CMS, "radiusDictionary.dat"))


def _event(n=3, **kw):
    from trax_amd import pbio
    rng = np.random.default_rng(5)
    cols = dict(x=rng.normal(size=n).astype(np.float32), y=rng.normal(size=n).astype(np.float32),
                z=rng.normal(size=n).astype(np.float32), detector_type=np.array([-1, 0, 1][:n], np.int8),
                layer=np.array([1, 2, 13][:n], np.uint64), detector_id=np.array([0, 150, 2**63 + 7][:n], np.uint64),
                hit_id=np.array([0, 1, 2**64 - 1][:n], np.uint64), sim_track_id=np.array([0, 5, 2**40][:n], np.uint64),
                sim_track_pt=np.array([0.0, 1.5, 20.25][:n], np.float32),
                has_sim_track_id=np.array([False, True, True][:n]), has_sim_track_pt=np.array([False, True, True][:n]))
    cols.update(kw)
    return pbio.PEvent(run=1, lumi=2**35, event=150, **cols)


def test_wire_format_known_answers():
    """protobuf's encoding guide: uint 150 in field 1 is 08 96 01; an enum of
    -1 is sign-extended to a ten-byte varint."""
    from trax_amd import pbio
    e = _event(1, detector_type=np.array([-1], np.int8))
    b = pbio.encode_event(e)
    assert b[:2] == bytes([0x08, 0x01])  # runNumber 1
    assert bytes([0x18, 0x96, 0x01]) in b  # eventNumber 150 (field 3)
    assert bytes([0x10] + [0xFF] * 9 + [0x01]) in b  # detectorType -1
    out = bytearray()
    pbio._put_varint(out, 150)
    assert bytes(out) == bytes([0x96, 0x01])
    assert pbio._varint(bytes([0x96, 0x01]), 0) == (150, 2)


def test_synthetic_round_trip_keeps_absent_fields(tmp_path):
    from trax_amd import pbio
    e = _event()
    p = str(tmp_path / "c.pb")
    pbio.write_event_container(p, [e, e])
    back = pbio.read_event_container(p)
    assert len(back) == 2
    for f in ("x", "y", "z", "detector_type", "layer", "detector_id", "hit_id", "sim_track_id",
              "sim_track_pt", "has_sim_track_id", "has_sim_track_pt"):
        assert np.array_equal(getattr(back[1], f), getattr(e, f)), f
    assert (back[0].run, back[0].lumi, back[0].event) == (1, 2**35, 150)
    assert pbio.encode_event(back[0]) == pbio.encode_event(e)


def test_container_re_encodes_byte_for_byte(events, tmp_path):
    from trax_amd import pbio
    p = str(tmp_path / "again.pb")
    pbio.write_event_container(p, events)
    assert open(p, "rb").read() == open(PB, "rb").read()


def test_container_contents(events):
    """10 CMS ttbar events of 941-2138 hits; barrel layers 1-5, 8, 9; hits
    of simulated-track 0 (noise) carry no simTrackPt."""
    assert len(events) == 10
    n = [e.n_hits for e in events]
    assert min(n) == 941 and max(n) == 2138 and sum(n) == 15252
    layers = set(int(v) for e in events for v in e.layer)
    assert layers == {1, 2, 3, 4, 5, 8, 9}
    for e in events:
        assert np.all(e.detector_type == 0)
        assert np.all(e.has_sim_track_id)
        noise = e.sim_track_id == 0
        assert not np.any(e.has_sim_track_pt[noise])
        assert np.all(e.sim_track_pt[~e.has_sim_track_pt] == 0)


def test_hits_agree_with_the_geometry(events, geom):
    """Detector ids resolve, to a module of the hit's own layer, and the
    hit radius lies within 0.6 cm of that layer's module-radius band."""
    for e in events:
        idx = np.array([geom.resolve(int(d)) for d in e.detector_id])
        assert np.all(idx >= 0)
        assert np.array_equal(geom.det_layer[idx], e.layer.astype(np.uint32))
        r = np.hypot(e.x.astype(np.float64), e.y.astype(np.float64))
        lo = np.array(geom.min_radius)[e.layer.astype(int) - 1]
        hi = np.array(geom.max_radius)[e.layer.astype(int) - 1]
        assert np.all(r > lo - 0.6) and np.all(r < hi + 0.6)


def test_geometry_radii_and_lookup(geom):
    ids, entry, layer = np.loadtxt(os.path.join(CMS, "detectorRadius.dat"), delimiter=",",
                                   dtype=np.int64, unpack=True)
    radius = np.loadtxt(os.path.join(CMS, "radiusDictionary.dat"), delimiter=",")[:, 1].astype(np.float32)
    for l in range(1, 14):
        rr = radius[entry[layer == l]]
        assert geom.min_radius[l - 1] == float(rr.min()) and geom.max_radius[l - 1] == float(rr.max())
    # pixel barrel layers 1-3 at r = 4.1-4.7, 7.0-7.6, 9.9-10.4 cm
    for l, (a, b) in enumerate([(4.1, 4.7), (7.0, 7.6), (9.9, 10.4)], start=1):
        assert a <= geom.min_radius[l - 1] < geom.max_radius[l - 1] <= b
    sample = np.random.default_rng(1).choice(len(ids), 200, replace=False)
    assert all(geom.resolve(int(ids[i])) == i for i in sample)
    assert geom.resolve(int(ids[0]) - 1) == -1 and geom.resolve(int(ids[-1]) + 1) == -1
    assert geom.resolve(int(ids[10]) + 1) == (-1 if ids[11] != ids[10] + 1 else 11)


def test_event_store_and_loaders(events, tmp_path):
    from trax_amd import pbio
    s = str(tmp_path / "store.pb")
    pbio.write_event_store(s, events)
    raw = open(s, "rb").read()
    assert raw[:4] == (10).to_bytes(4, "little") and raw[4] == 1  # count, then entry type Event
    back = pbio.read_event_store(s)
    assert [pbio.encode_event(e) for e in back] == [pbio.encode_event(e) for e in events]
    assert [e.event for e in pbio.load(s, "store", skip=2, max_events=5)] == [e.event for e in events[2:7]]
    assert [e.event for e in pbio.load(PB, "standard", skip=8, max_events=5)] == [e.event for e in events[8:]]
    assert len(pbio.load(PB)) == 10
    rep = pbio.load(PB, "repeated", skip=3, max_events=7)
    assert len(rep) == 4 and all(e is rep[0] for e in rep) and rep[0].event == events[3].event
    conv = pbio.store_to_container(back, skip=1, max_events=3, multiplicity=2)
    assert [e.event for e in conv] == [events[1].event] * 2 + [events[2].event] * 2
    with pytest.raises(ValueError):
        pbio.load(PB, "nope")
    with pytest.raises(ValueError):
        pbio.load(PB, "repeated", max_events=-1)


def _expected_order(e, used, max_layer, min_pt, only_tracks, max_tracks=-1):
    """(hit indexes in the chain's order, findable track ids) by sorting keys:
    layer, then (non-findable before findable), then track id, then file
    order."""
    tid = (e.sim_track_id & np.uint64(0xFFFFFFFF)).astype(np.int64)
    keep = e.sim_track_pt >= np.float32(min_pt)
    if only_tracks:
        keep &= e.sim_track_id != 0
    lay = e.layer.astype(np.int64)
    ids = np.unique(tid[keep])
    findable, ok_ids = [], []
    for t in ids:  # ascending, like the reference's std::map
        if len(findable) == max_tracks:
            break
        if only_tracks and t == 0:
            continue
        cov = set(lay[keep & (tid == t)].tolist())
        if set(used) <= cov:
            findable.append(t)
        ok_ids.append(t)
    is_f = np.isin(tid, findable)
    sel = keep & np.isin(tid, ok_ids) & (lay <= max_layer)
    if only_tracks:
        sel &= is_f
    idx = np.nonzero(sel)[0]
    order = np.lexsort((idx, tid[idx], is_f[idx].astype(np.int64), lay[idx]))
    return idx[order], findable


@pytest.mark.parametrize("only_tracks,min_pt,max_tracks", [(True, 1.0, -1), (False, 1.0, -1), (True, 0.0, -1),
                                                          (False, 0.0, -1), (True, 1.0, 5)])
def test_load_events_order_and_selection(events, geom, only_tracks, min_pt, max_tracks):
    from trax_amd import cmsdata
    trip = cmsdata.reference_triplets()
    used = sorted({l for t in trip for l in (t.layer1, t.layer2, t.layer3)})
    assert used == [1, 2, 3, 4, 5, 8]
    batch, tracks = cmsdata.load_events(events, geom, trip, min_pt, max_tracks, only_tracks)
    L = 8
    assert batch.n_layers == L and batch.n_events == 10
    for k, e in enumerate(events):
        want, findable = _expected_order(e, used, L, min_pt, only_tracks, max_tracks)
        assert sorted(tracks[k]) == [int(t) for t in findable]
        a = int(batch.event_offset[k])
        b = int(batch.event_offset[k + 1]) if k + 1 < 10 else batch.n_hits
        assert b - a == len(want)
        assert np.array_equal(batch.x[a:b], e.x[want]) and np.array_equal(batch.z[a:b], e.z[want])
        assert np.array_equal(batch.layer[a:b], e.layer[want].astype(np.uint32))
        assert np.array_equal(batch.hit_id[a:b], (e.sim_track_id[want] & np.uint64(0xFFFFFFFF)).astype(np.uint32))
        assert np.array_equal(geom.det_ids[batch.det_id[a:b]], e.detector_id[want])
        assert np.all(batch.evt[a:b] == k)
        n = batch.layer_nhits[k * L:(k + 1) * L].astype(np.int64)
        off = batch.layer_offset[k * L:(k + 1) * L].astype(np.int64)
        assert np.array_equal(n, np.bincount(e.layer[want].astype(np.int64) - 1, minlength=L))
        assert np.array_equal(off, np.concatenate([[0], np.cumsum(n)[:-1]]))
        for t, hits in tracks[k].items():
            assert len(hits) >= len(used) and all(int(e.sim_track_id[i]) & 0xFFFFFFFF == t for i in hits)


def test_ttbar_workload_runs_through_the_oracle(events, geom):
    """The reference's ttbar configuration (grid 50 x 8, minPt 1 GeV, four
    layer triplets, eight layers) on the real events: the oracle chain runs
    with nothing left unwritten, and findable tracks reappear as valid
    triplets (three hits of one simulated track)."""
    from oracle import bridge
    from trax_amd import cmsdata
    batch, tracks = cmsdata.load_events(events, geom, cmsdata.reference_triplets(), 1.0)
    sc = cmsdata.ttbar_workload(geom)
    sc.validate()
    o = bridge.run_chain(sc, batch)
    assert o["n_pairs_unwritten"] == 0
    t = o["tracklets"].reshape(-1, 3)
    hid = o["hit_id"]  # tracklets index the grid-sorted hits
    valid = (hid[t[:, 0]] == hid[t[:, 1]]) & (hid[t[:, 0]] == hid[t[:, 2]])
    assert valid.sum() > 0.5 * sum(len(x) for x in tracks)
