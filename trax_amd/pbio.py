"""Reader and writer of the reference's event files: a protobuf
`PEventContainer` (src/datastructures/serialize/Event.proto:1-37), decoded
from the wire format directly (no protobuf runtime needed).

    PEventContainer { repeated PEvent events = 1; }
    PEvent { uint64 runNumber = 1; uint64 lumiSection = 2; uint64 eventNumber = 3;
             repeated PHit hits = 4; }
    PHit { PGlobalPoint position = 1; DetectorType detectorType = 2 (enum, -1..1);
           uint64 layer = 3; uint64 detectorId = 4; uint64 hitId = 5;
           optional uint64 simTrackID = 6; optional float simTrackPt = 7; }
    PGlobalPoint { float x = 1; float y = 2; float z = 3; }

The reference reads the whole container with ParseFromCodedStream
(src/triplet_builder/EventLoader.h:19-44); each event comes back here as a
`PEvent` of numpy columns, one entry per hit in file order, with the presence
of the two optional fields kept so that `write_event_container` re-encodes a
file byte for byte.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np


@dataclass
class PEvent:
    run: int
    lumi: int
    event: int
    x: np.ndarray  # float32
    y: np.ndarray
    z: np.ndarray
    detector_type: np.ndarray  # int8: -1 backward, 0 barrel, 1 forward
    layer: np.ndarray  # uint64
    detector_id: np.ndarray
    hit_id: np.ndarray
    sim_track_id: np.ndarray  # uint64, 0 when absent
    sim_track_pt: np.ndarray  # float32, 0 when absent
    has_sim_track_id: np.ndarray  # bool: the optional field is present
    has_sim_track_pt: np.ndarray

    @property
    def n_hits(self) -> int:
        return int(self.x.shape[0])


def _varint(buf: bytes, pos: int) -> Tuple[int, int]:
    v = shift = 0
    while True:
        b = buf[pos]
        pos += 1
        v |= (b & 0x7F) << shift
        if b < 0x80:
            return v, pos
        shift += 7
        if shift > 63:
            raise ValueError("varint longer than 10 bytes")


def _fields(buf: bytes, pos: int, end: int):
    """(field number, wire type, value) of a message in buf[pos:end]; LEN
    values are (start, stop) slices, fixed32 values raw bytes."""
    while pos < end:
        key, pos = _varint(buf, pos)
        num, wt = key >> 3, key & 7
        if wt == 0:
            v, pos = _varint(buf, pos)
        elif wt == 2:
            n, pos = _varint(buf, pos)
            v = (pos, pos + n)
            pos += n
        elif wt == 5:
            v = buf[pos:pos + 4]
            pos += 4
        elif wt == 1:
            v = buf[pos:pos + 8]
            pos += 8
        else:
            raise ValueError(f"unsupported wire type {wt}")
        if pos > end:
            raise ValueError("truncated message")
        yield num, wt, v


def _signed(v: int) -> int:  # int32 enum encoded as a 64-bit varint
    return v - (1 << 64) if v >= 1 << 63 else v


def _event(buf: bytes, start: int, stop: int) -> PEvent:
    run = lumi = num = 0
    cols = {k: [] for k in ("x", "y", "z", "t", "l", "d", "h", "s", "p", "hs", "hp")}
    for f, wt, v in _fields(buf, start, stop):
        if f == 1 and wt == 0:
            run = v
        elif f == 2 and wt == 0:
            lumi = v
        elif f == 3 and wt == 0:
            num = v
        elif f == 4 and wt == 2:
            x = y = z = 0.0
            t = l = d = h = s = 0
            p = 0.0
            hs = hp = False
            for g, gt, u in _fields(buf, v[0], v[1]):
                if g == 1 and gt == 2:
                    for q, qt, w in _fields(buf, u[0], u[1]):
                        if qt == 5:
                            val = struct.unpack("<f", w)[0]
                            if q == 1:
                                x = val
                            elif q == 2:
                                y = val
                            elif q == 3:
                                z = val
                elif g == 2 and gt == 0:
                    t = _signed(u)
                elif g == 3 and gt == 0:
                    l = u
                elif g == 4 and gt == 0:
                    d = u
                elif g == 5 and gt == 0:
                    h = u
                elif g == 6 and gt == 0:
                    s, hs = u, True
                elif g == 7 and gt == 5:
                    p, hp = struct.unpack("<f", u)[0], True
            for k, val in zip(("x", "y", "z", "t", "l", "d", "h", "s", "p", "hs", "hp"),
                              (x, y, z, t, l, d, h, s, p, hs, hp)):
                cols[k].append(val)
    f32 = lambda k: np.asarray(cols[k], np.float32)
    u64 = lambda k: np.asarray(cols[k], np.uint64)
    return PEvent(run, lumi, num, f32("x"), f32("y"), f32("z"), np.asarray(cols["t"], np.int8),
                  u64("l"), u64("d"), u64("h"), u64("s"), f32("p"), np.asarray(cols["hs"], bool),
                  np.asarray(cols["hp"], bool))


def read_event_container(path: str) -> List[PEvent]:
    """All events of a PEventContainer file, in file order."""
    with open(path, "rb") as f:
        buf = f.read()
    return [_event(buf, v[0], v[1]) for fnum, wt, v in _fields(buf, 0, len(buf)) if fnum == 1 and wt == 2]


def _put_varint(out: bytearray, v: int) -> None:
    v &= (1 << 64) - 1
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return


def _put_len(out: bytearray, field: int, payload: bytes) -> None:
    _put_varint(out, (field << 3) | 2)
    _put_varint(out, len(payload))
    out += payload


def encode_event(e: PEvent) -> bytes:
    out = bytearray()
    for fnum, v in ((1, e.run), (2, e.lumi), (3, e.event)):
        _put_varint(out, fnum << 3)
        _put_varint(out, int(v))
    for i in range(e.n_hits):
        pt = bytearray()
        for fnum, v in ((1, e.x[i]), (2, e.y[i]), (3, e.z[i])):
            _put_varint(pt, (fnum << 3) | 5)
            pt += struct.pack("<f", float(v))
        hit = bytearray()
        _put_len(hit, 1, bytes(pt))
        for fnum, v in ((2, int(e.detector_type[i])), (3, int(e.layer[i])), (4, int(e.detector_id[i])),
                        (5, int(e.hit_id[i]))):
            _put_varint(hit, fnum << 3)
            _put_varint(hit, v)
        if e.has_sim_track_id[i]:
            _put_varint(hit, 6 << 3)
            _put_varint(hit, int(e.sim_track_id[i]))
        if e.has_sim_track_pt[i]:
            _put_varint(hit, (7 << 3) | 5)
            hit += struct.pack("<f", float(e.sim_track_pt[i]))
        _put_len(out, 4, bytes(hit))
    return bytes(out)


def write_event_container(path: str, events: List[PEvent]) -> None:
    out = bytearray()
    for e in events:
        _put_len(out, 1, encode_event(e))
    with open(path, "wb") as f:
        f.write(out)


# EventStore files (src/datastructures/serialize/PEventStore.h:25-245): a
# little-endian uint32 event count, then per event the varint entry type
# (1 = Event), the varint payload size and the serialized PEvent.
_ENTRY_EVENT = 1


def read_event_store(path: str) -> List[PEvent]:
    """All events of an EventStore file (GenericInputStore::readNextElement,
    PEventStore.h:163-199: reading stops at the stored count, at an entry that
    is not an event or at an empty payload)."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 4:
        raise ValueError("EventStore file shorter than its header")
    n = struct.unpack_from("<I", buf, 0)[0]
    pos, out = 4, []
    while len(out) < n and pos < len(buf):
        etype, pos = _varint(buf, pos)
        if etype != _ENTRY_EVENT:
            break
        size, pos = _varint(buf, pos)
        if size == 0 or pos + size > len(buf):
            break
        out.append(_event(buf, pos, pos + size))
        pos += size
    return out


def write_event_store(path: str, events: List[PEvent]) -> None:
    """GenericOutputStore::storeElement for every event, then the count in
    the header (PEventStore.h:70-115)."""
    out = bytearray(struct.pack("<I", len(events)))
    for e in events:
        payload = encode_event(e)
        _put_varint(out, _ENTRY_EVENT)
        _put_varint(out, len(payload))
        out += payload
    with open(path, "wb") as f:
        f.write(out)


def store_to_container(events: List[PEvent], skip: int = 0, max_events: int = -1,
                       multiplicity: int = 1) -> List[PEvent]:
    """evtStoreToContainer (src/triplet_builder/evtStoreToContainer.cpp:55-75):
    events [skip, last) of a store, each repeated `multiplicity` times, where
    last = all events, or min(max_events, count) (an absolute index)."""
    last = len(events) if max_events == -1 else min(max_events, len(events))
    return [e for e in events[skip:last] for _ in range(multiplicity)]


def load(path: str, loader: str = "standard", skip: int = 0, max_events: int = -1) -> List[PEvent]:
    """The events one run processes, as EventLoaderFactory's three loaders and
    the driver's event range give them (EventLoader.h:15-168,
    TripletBuilder.cpp:155-160):

    * "standard": a PEventContainer file; events [skip, last), last = all
      events or min(skip + max_events, count);
    * "repeated": event `skip` of a PEventContainer file, repeated; the
      loader reports max_events events and the driver's range starts at skip,
      so a run sees max_events - skip copies;
    * "store": an EventStore file, with the standard range.
    """
    if loader == "repeated":
        if max_events < 0:
            raise ValueError("the repeated loader needs max_events >= 0")
        evs = read_event_container(path)
        if not 0 <= skip < len(evs):
            raise ValueError(f"skip {skip} outside the {len(evs)} events of {path}")
        return [evs[skip]] * max(0, max_events - skip)
    if loader == "standard":
        evs = read_event_container(path)
    elif loader == "store":
        evs = read_event_store(path)
    else:
        raise ValueError(f"unknown event loader {loader!r} (standard, repeated, store)")
    last = len(evs) if max_events == -1 else min(skip + max_events, len(evs))
    return evs[skip:last]
