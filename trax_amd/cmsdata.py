"""Real-data input: the reference's CMS event files and detector geometry
turned into the hit layout of the seeding chain.

* `Geometry` reads configs/detectorRadius.dat (detector id, radius
  dictionary entry, layer) and configs/radiusDictionary.dat (entry, radius)
  the way src/triplet_builder/TripletBuilder.cpp:100-141 does: per layer 1..13
  the smallest and largest module radius (GeometrySupplement), and the
  detector-id -> index lookup of DetectorGeometry::resolveDetId
  (src/datastructures/DetectorGeometry.cpp:4-30, a binary search in the
  id-sorted list).
* `load_events` restates HitCollection::addEvent
  (src/datastructures/HitCollection.cpp:20-146): hits grouped by simulated
  track, hits below minPt dropped, a track kept as "findable" only when it
  has a hit in every layer the layer triplets use, every event's hits bucketed
  by layer (hits of non-findable tracks first when they are kept, then the
  findable tracks' hits, tracks in ascending id order), hitId = the
  simulated-track id, detId = the resolved index.
* `reference_triplets` is the layer-triplet table of configs/layerTriplets.xml
  and `ttbar_workload` the configuration of configs/ttBar.reco.conf (grid 50
  z x 8 phi, minPt 1 GeV, only complete tracks).

Where the reference is unspecified this module fixes: a track's hits in one
layer keep their file order (the reference sorts with std::sort, which may
reorder equal layers); hits on layers above the largest triplet layer do not
mark layers as covered (the reference writes them past the end of its
coverage vector).
"""
from __future__ import annotations

import csv
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import events as ev
from . import pbio


@dataclass
class Geometry:
    det_ids: np.ndarray  # uint64, ascending (file order)
    det_layer: np.ndarray  # uint32
    min_radius: List[float]  # layers 1..13 (index layer - 1), float32 values
    max_radius: List[float]

    @staticmethod
    def load(detector_radius_dat: str, radius_dictionary_dat: str, n_layers: int = 13) -> "Geometry":
        with open(radius_dictionary_dat) as f:
            radius = [np.float32(float(r[1])) for r in csv.reader(f) if r]
        ids, layers, entries = [], [], []
        with open(detector_radius_dat) as f:
            for r in csv.reader(f):
                if not r:
                    continue
                ids.append(int(r[0]))
                entries.append(int(r[1]))
                layers.append(int(r[2]))
        lo = {l: np.float32(10000.0) for l in range(1, n_layers + 1)}
        hi = {l: np.float32(0.0) for l in range(1, n_layers + 1)}
        for e, l in zip(entries, layers):
            rr = radius[e]
            if rr > hi[l]:
                hi[l] = rr
            if rr < lo[l]:
                lo[l] = rr
        return Geometry(np.asarray(ids, np.uint64), np.asarray(layers, np.uint32),
                        [float(lo[l]) for l in range(1, n_layers + 1)],
                        [float(hi[l]) for l in range(1, n_layers + 1)])

    def resolve(self, det_id: int) -> int:
        """Index of det_id in the id-sorted detector list, or -1."""
        i = int(np.searchsorted(self.det_ids, np.uint64(det_id)))
        return i if i < self.det_ids.shape[0] and int(self.det_ids[i]) == det_id else -1


def reference_triplets() -> List[ev.LayerTriplet]:
    """configs/layerTriplets.xml of the reference, row by row."""
    return [
        ev.LayerTriplet(1, 2, 3, dtheta_cut=0.01, sigma_z=0.037, dphi_cut=0.075, sigma_phi=0.032,
                        tip_cut=0.5, pair_spread_z=2, z0=10, pair_spread_phi=2, n_candidates=4),
        ev.LayerTriplet(2, 3, 4, dtheta_cut=0.2, sigma_z=0.207, dphi_cut=0.2, sigma_phi=0.102,
                        tip_cut=0.5, pair_spread_z=2, z0=10, pair_spread_phi=2, n_candidates=16),
        ev.LayerTriplet(3, 4, 5, dtheta_cut=0.4, sigma_z=0.207, dphi_cut=0.4, sigma_phi=0.102,
                        tip_cut=1.0, pair_spread_z=3, z0=15, pair_spread_phi=2, n_candidates=64),
        ev.LayerTriplet(4, 5, 8, dtheta_cut=0.5, sigma_z=0.207, dphi_cut=0.4, sigma_phi=0.102,
                        tip_cut=3.5, pair_spread_z=4, z0=30, pair_spread_phi=2, n_candidates=128),
    ]


TrackList = Dict[int, List[int]]  # simulated-track id -> hit indexes of the event (file order)


def _add_event(e: pbio.PEvent, geom: Geometry, triplets: Sequence[ev.LayerTriplet], min_pt: float,
               max_tracks: int, only_tracks: bool) -> Tuple[List[np.ndarray], TrackList]:
    """One event's hits bucketed by layer 1..maxLayer (HitCollection.cpp:20-104)
    and its findable tracks."""
    used = sorted({l for t in triplets for l in (t.layer1, t.layer2, t.layer3)})
    max_layer = max(t.layer3 for t in triplets)
    pt = e.sim_track_pt
    tid = e.sim_track_id
    tracks: TrackList = {}
    for i in range(e.n_hits):
        if only_tracks and int(tid[i]) == 0:
            continue
        if pt[i] < np.float32(min_pt):
            continue
        tracks.setdefault(int(tid[i]) & 0xFFFFFFFF, []).append(i)  # tTrackList keys are uint
    layer = e.layer.astype(np.int64)
    layers: Dict[int, List[int]] = {}
    findable: TrackList = {}
    found = 0
    for t in sorted(tracks):
        if found == max_tracks:
            break
        if only_tracks and t == 0:
            continue
        hits = sorted(tracks[t], key=lambda i: layer[i])  # stable: file order within a layer
        covered = {int(layer[i]) for i in hits if int(layer[i]) <= max_layer}
        missing = any(l not in covered for l in used)
        if pt[hits[0]] < np.float32(min_pt):
            continue
        for i in hits:
            if not missing:
                findable.setdefault(t, []).append(i)
            elif not only_tracks:
                layers.setdefault(int(layer[i]), []).append(i)
        if not missing:
            found += 1
    for t in sorted(findable):
        for i in findable[t]:
            if layer[i] <= max_layer:
                layers.setdefault(int(layer[i]), []).append(i)
    return [np.asarray(layers.get(l, []), np.int64) for l in range(1, max_layer + 1)], findable


def load_events(events: Sequence[pbio.PEvent], geom: Geometry, triplets: Sequence[ev.LayerTriplet],
                min_pt: float, max_tracks: int = -1, only_tracks: bool = True
                ) -> Tuple[ev.EventBatch, List[TrackList]]:
    """Events in the seeding chain's layout (event-major, layers ascending),
    and per event its findable tracks (hit indexes into that event's file
    order)."""
    L = max(t.layer3 for t in triplets)
    cols = {k: [] for k in ("x", "y", "z", "layer", "det_id", "hit_id", "evt")}
    ev_off, l_n, l_off = [], [], []
    all_tracks = []
    total = 0
    for k, e in enumerate(events):
        per_layer, findable = _add_event(e, geom, triplets, min_pt, max_tracks, only_tracks)
        all_tracks.append(findable)
        ev_off.append(total)
        rel = 0
        for l, idx in enumerate(per_layer, start=1):
            l_n.append(len(idx))
            l_off.append(rel)
            rel += len(idx)
            cols["x"].append(e.x[idx])
            cols["y"].append(e.y[idx])
            cols["z"].append(e.z[idx])
            cols["layer"].append(e.layer[idx].astype(np.uint32))
            det = [geom.resolve(int(d) & 0xFFFFFFFF) for d in e.detector_id[idx]]
            det = [d if d >= 0 else int(raw) & 0xFFFFFFFF for d, raw in zip(det, e.detector_id[idx])]
            cols["det_id"].append(np.asarray(det, np.uint32))
            cols["hit_id"].append((e.sim_track_id[idx] & np.uint64(0xFFFFFFFF)).astype(np.uint32))
            cols["evt"].append(np.full(len(idx), k, np.uint32))
        total += rel
    cat = {k: (np.concatenate(v) if v else np.zeros(0)) for k, v in cols.items()}
    batch = ev.EventBatch(cat["x"].astype(np.float32), cat["y"].astype(np.float32),
                          cat["z"].astype(np.float32), cat["layer"].astype(np.uint32),
                          cat["det_id"].astype(np.uint32), cat["hit_id"].astype(np.uint32),
                          cat["evt"].astype(np.uint32), np.asarray(ev_off, np.uint32),
                          np.asarray(l_n, np.uint32), np.asarray(l_off, np.uint32), L)
    return batch, all_tracks


def ttbar_workload(geom: Geometry, triplets: Optional[Sequence[ev.LayerTriplet]] = None,
                   min_pt: float = 1.0) -> ev.SeedConfig:
    """configs/ttBar.reco.conf: grid 50 z x 8 phi, minPt 1 GeV, 256 threads,
    the layer-triplet table of configs/layerTriplets.xml."""
    trip = list(triplets) if triplets is not None else reference_triplets()
    L = max(t.layer3 for t in trip)
    return ev.SeedConfig(n_z=50, n_phi=8, min_radius=geom.min_radius[:L], max_radius=geom.max_radius[:L],
                         triplets=trip, min_pt=min_pt, threads=256)
