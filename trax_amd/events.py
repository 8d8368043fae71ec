"""Seeded synthetic events and seeding configurations.

Events are helices from a vertex on the beam line in a uniform solenoid field,
intersected with barrel cylinders, smeared, and stored in the reference's hit
layout: SoA columns (x, y, z: float32; layer, detId, hitId, evt: uint32),
event-major, layers ascending inside an event, with per-event and
per-(event, layer) hit counts and offsets (reference: src/datastructures/
HitCollection.cpp:106-146, EventSupplement.h, LayerSupplement.h).  `hitId`
carries the simulated track id, 0 for noise, as in HitCollection.cpp:136-138.

Every event is generated from its own counter-based substream
(seed, config id, event index), so an event has the same hits whether it is
generated alone or inside a batch.

The five workloads follow the five configurations the seeding path is
measured on (plain single muons, pile-up 50, uneven batched pile-up 20-200,
pile-up 200 on six layers, and a serving mix).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

SEED_BASE = 0x7472617800  # "trax" + 00

MIN_Z = np.float32(-300.0)  # Grid.h:33-37
MAX_Z = np.float32(300.0)
MIN_PHI = np.float32(-math.pi)
MAX_PHI = np.float32(math.pi)
IDEAL_BZ = 3.8112  # TripletConfiguration.h:70
B_FIELD_T = 3.8  # generator field


def min_radius_curvature(min_pt: float, bz: float = IDEAL_BZ) -> np.float32:
    """TripletConfiguration.cpp:145-152: minPt and Bz enter as float, the
    formula runs in double, the result is returned as float."""
    q = 1.602177e-19
    gev_c = 5.344286e-19
    pt = float(np.float32(min_pt))
    b = float(np.float32(bz))
    return np.float32(pt * gev_c / (b * q) * 1e2)


@dataclass
class LayerTriplet:
    """One row of the layer-triplet table (TripletConfiguration.h:9-33)."""

    layer1: int
    layer2: int
    layer3: int
    dtheta_cut: float = 0.01
    sigma_z: float = 0.1
    dphi_cut: float = 0.1
    sigma_phi: float = 0.1
    tip_cut: float = 0.5
    pair_spread_z: int = 2
    z0: int = 15  # parsed as uint by the reference (TripletConfiguration.cpp:91)
    pair_spread_phi: int = 2
    n_candidates: int = 4


@dataclass
class SeedConfig:
    """Grid binning, per-layer radius bands and the layer-triplet table."""

    n_z: int
    n_phi: int
    min_radius: List[float]  # index: layer - 1
    max_radius: List[float]
    triplets: List[LayerTriplet]
    min_pt: float = 1.0
    threads: int = 256  # work-group size of the reference (orders pairs only)
    # pair generator: "beamspot" (PairGeneratorBeamspot) or "sector"
    # (PairGeneratorSector, pairSpreadZ x pairSpreadPhi cells, no cut)
    pair_generator: str = "beamspot"

    @property
    def n_layers(self) -> int:
        return max(t.layer3 for t in self.triplets)

    @property
    def sector_size_z(self) -> np.float32:  # Grid.h:9-11 (float arithmetic)
        return np.float32((MAX_Z - MIN_Z) / np.float32(self.n_z))

    @property
    def sector_size_phi(self) -> np.float32:  # Grid.h:13-15
        return np.float32((MAX_PHI - MIN_PHI) / np.float32(self.n_phi))

    @property
    def min_radius_curvature(self) -> np.float32:
        return min_radius_curvature(self.min_pt)

    @property
    def cells_per_layer(self) -> int:
        return (self.n_z + 1) * (self.n_phi + 1)

    def validate(self) -> None:
        for p, t in enumerate(self.triplets):
            # Predictor and filter take the layer triplet as innermost layer - 1
            # (TripletThetaPhiPredictor.h:79, TripletThetaPhiFilter.h:73).
            if t.layer1 != p + 1:
                raise ValueError(f"layer triplet {p} must have layer1 == {p + 1}")


@dataclass
class EventBatch:
    """A group of events in the reference's hit layout (host arrays)."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    layer: np.ndarray
    det_id: np.ndarray
    hit_id: np.ndarray
    evt: np.ndarray
    event_offset: np.ndarray  # [E]
    layer_nhits: np.ndarray  # [E * L]
    layer_offset: np.ndarray  # [E * L], relative to the event start
    n_layers: int

    @property
    def n_events(self) -> int:
        return int(self.event_offset.shape[0])

    @property
    def n_hits(self) -> int:
        return int(self.x.shape[0])

    def event(self, e: int) -> "EventBatch":
        return concat_events([self], subset=[e])

    def columns(self):
        return (self.x, self.y, self.z, self.layer, self.det_id, self.hit_id, self.evt)


@dataclass
class Workload:
    name: str
    config_id: int
    seed_config: SeedConfig
    n_events: int
    radii: List[float]
    half_length: float = 26.6
    pileup: Optional[Sequence[int]] = None  # (lo, hi) inclusive, or fixed
    tracks_per_vertex: float = 30.0
    pt_mean: float = 0.7
    pt_cut: float = 0.3
    eta_max: float = 2.5
    noise_frac: float = 0.0
    zv_sigma: float = 5.0
    muons: int = 0  # > 0: config-0 style single muons


def _band(radii: Sequence[float], half: float = 0.25):
    return [r - half for r in radii], [r + half for r in radii]


def config_workload(config_id: int, n_events: Optional[int] = None) -> Workload:
    """The five measured workloads, by configuration number 0..4."""
    if config_id == 0:
        radii = [4.4, 7.3, 10.2]
        lo, hi = _band(radii)
        # the reference's own first layer-triplet row (configs/layerTriplets.xml:5-19)
        t = LayerTriplet(1, 2, 3, dtheta_cut=0.01, sigma_z=0.037, dphi_cut=0.075,
                         sigma_phi=0.032, tip_cut=0.5, pair_spread_z=2, z0=10,
                         pair_spread_phi=2, n_candidates=4)
        sc = SeedConfig(n_z=16, n_phi=32, min_radius=lo, max_radius=hi, triplets=[t], min_pt=1.0)
        return Workload("plain-10mu-3layer", 0, sc, n_events or 1, radii, muons=10, eta_max=1.0,
                        zv_sigma=0.0)
    if config_id == 1:
        radii = [4.4, 7.3, 10.2]
        lo, hi = _band(radii)
        sc = SeedConfig(n_z=64, n_phi=128, min_radius=lo, max_radius=hi,
                        triplets=[LayerTriplet(1, 2, 3)], min_pt=0.3)
        return Workload("plain-pu50-3layer", 1, sc, n_events or 1, radii, pileup=(50, 50),
                        tracks_per_vertex=30, pt_cut=0.3, noise_frac=0.02)
    if config_id == 2:
        radii = [4.4, 7.3, 10.2, 16.0]
        lo, hi = _band(radii)
        # 256 phi x 128 z: the autotuner's pick over 64-256 phi x 32-128 z for
        # this batch (chain 2.36 ms against 3.02 ms at 128 x 64)
        sc = SeedConfig(n_z=128, n_phi=256, min_radius=lo, max_radius=hi,
                        triplets=[LayerTriplet(1, 2, 3), LayerTriplet(2, 3, 4)], min_pt=0.3)
        return Workload("batched-pu20to200-4layer", 2, sc, n_events or 256, radii,
                        pileup=(20, 200), tracks_per_vertex=30, pt_cut=0.3, noise_frac=0.02)
    if config_id == 3:
        radii = [4.4, 7.3, 10.2, 25.5, 33.9, 41.8]
        lo, hi = _band(radii)
        trip = [LayerTriplet(1, 2, 3), LayerTriplet(2, 3, 4), LayerTriplet(3, 4, 5),
                LayerTriplet(4, 5, 6)]
        # 256 z bins: the binning the autotuner (trax_amd.autotune) picks for this
        # workload; it yields the same pairs, candidates and triplets as 64 z bins
        # (the barrel spans |z| < 26.6 cm of the grid's fixed +-300 cm, so finer
        # z bins mainly shorten the window walks)
        sc = SeedConfig(n_z=256, n_phi=128, min_radius=lo, max_radius=hi, triplets=trip, min_pt=0.5)
        return Workload("batched-pu200-6layer", 3, sc, n_events or 64, radii, pileup=(200, 200),
                        tracks_per_vertex=40, pt_cut=0.5, noise_frac=0.05)
    if config_id == 4:
        w = config_workload(3, n_events or 512)
        w.name = "serving-pu20-pu200-mix"
        w.config_id = 4
        w.pileup = None  # per event: 70% PU20, 30% PU200 (event_pileup)
        return w
    raise ValueError(f"unknown configuration {config_id}")


def event_pileup(w: Workload, e: int, rng: np.random.Generator) -> int:
    if w.config_id == 4:
        return 20 if rng.random() < 0.7 else 200
    lo, hi = w.pileup
    return int(lo) if lo == hi else int(rng.integers(lo, hi + 1))


def _gen_tracks(w: Workload, rng: np.random.Generator):
    """Track parameters (pt, eta, phi0, charge, z vertex)."""
    if w.muons:
        n = w.muons
        pt = rng.uniform(1.0, 10.0, n)
        eta = rng.uniform(-w.eta_max, w.eta_max, n)
        phi0 = rng.uniform(0.0, 2 * math.pi, n)
        q = rng.choice(np.array([-1.0, 1.0]), n)
        zv = np.zeros(n)
        return pt, eta, phi0, q, zv
    pu = event_pileup(w, 0, rng)
    zvert = rng.normal(0.0, w.zv_sigma, pu)
    ntrk = rng.poisson(w.tracks_per_vertex, pu)
    zv = np.repeat(zvert, ntrk)
    n = zv.shape[0]
    pt = rng.exponential(w.pt_mean, n)
    eta = rng.uniform(-w.eta_max, w.eta_max, n)
    phi0 = rng.uniform(-math.pi, math.pi, n)
    q = rng.choice(np.array([-1.0, 1.0]), n)
    keep = pt > w.pt_cut
    return pt[keep], eta[keep], phi0[keep], q[keep], zv[keep]


def generate_event(w: Workload, e: int, seed: int = SEED_BASE) -> dict:
    """Hits of event e as per-layer lists (layer 1..L), noise first (id 0)."""
    rng = np.random.default_rng([seed, w.config_id, e])
    pt, eta, phi0, q, zv = _gen_tracks(w, rng)
    n_trk = pt.shape[0]
    R = pt / (0.3 * B_FIELD_T) * 100.0  # cm
    cot = np.sinh(eta)
    layers = []
    for li, r in enumerate(w.radii):
        ok = r < 2.0 * R
        s_arg = np.where(ok, r / (2.0 * np.where(ok, R, 1.0)), 0.0)
        alpha = np.arcsin(s_arg)
        s = 2.0 * R * alpha
        phi = phi0 - q * alpha + rng.normal(0.0, 1e-3, n_trk) / r  # 10 um r-phi
        z = zv + s * cot + rng.normal(0.0, 2e-3, n_trk)  # 20 um z
        keep = ok & (np.abs(z) < w.half_length)
        ids = np.nonzero(keep)[0]
        xs = (r * np.cos(phi[keep])).astype(np.float32)
        ys = (r * np.sin(phi[keep])).astype(np.float32)
        zs = z[keep].astype(np.float32)
        hid = (ids + 1).astype(np.uint32)
        n_noise = int(round(w.noise_frac * ids.shape[0])) if w.noise_frac > 0 else 0
        if n_noise:
            nphi = rng.uniform(-math.pi, math.pi, n_noise)
            nz = rng.uniform(-w.half_length, w.half_length, n_noise)
            xs = np.concatenate([(r * np.cos(nphi)).astype(np.float32), xs])
            ys = np.concatenate([(r * np.sin(nphi)).astype(np.float32), ys])
            zs = np.concatenate([nz.astype(np.float32), zs])
            hid = np.concatenate([np.zeros(n_noise, np.uint32), hid])
        phi_all = np.arctan2(ys.astype(np.float64), xs.astype(np.float64))
        n_mod_phi = 16 * (li + 1)
        mod = np.floor((phi_all + math.pi) / (2 * math.pi) * n_mod_phi).astype(np.int64) % n_mod_phi
        zmod = np.clip(np.floor((zs + w.half_length) / 6.65), 0, 7).astype(np.int64)
        det = ((li + 1) * 10000 + mod * 8 + zmod).astype(np.uint32)
        layers.append((xs, ys, zs, hid, det))
    return {"layers": layers}


def make_batch(w: Workload, events: Optional[Sequence[int]] = None, seed: int = SEED_BASE) -> EventBatch:
    """Concatenate events (by index) into one group in the reference layout."""
    if events is None:
        events = range(w.n_events)
    events = list(events)
    L = w.seed_config.n_layers
    cols = {k: [] for k in ("x", "y", "z", "layer", "det", "hid", "evt")}
    ev_off = np.zeros(len(events), np.uint32)
    l_n = np.zeros(len(events) * L, np.uint32)
    l_off = np.zeros(len(events) * L, np.uint32)
    total = 0
    for k, e in enumerate(events):
        ev = generate_event(w, e, seed)
        ev_off[k] = total
        off = 0
        for li in range(L):
            xs, ys, zs, hid, det = ev["layers"][li]
            n = xs.shape[0]
            l_n[k * L + li] = n
            l_off[k * L + li] = off
            off += n
            cols["x"].append(xs)
            cols["y"].append(ys)
            cols["z"].append(zs)
            cols["layer"].append(np.full(n, li + 1, np.uint32))
            cols["det"].append(det)
            cols["hid"].append(hid)
            cols["evt"].append(np.full(n, k, np.uint32))
        total += off

    def cat(key, dt):
        return np.ascontiguousarray(np.concatenate(cols[key]).astype(dt)) if cols[key] else np.zeros(0, dt)

    return EventBatch(cat("x", np.float32), cat("y", np.float32), cat("z", np.float32),
                      cat("layer", np.uint32), cat("det", np.uint32), cat("hid", np.uint32),
                      cat("evt", np.uint32), ev_off, l_n, l_off, L)


def concat_events(batches: Sequence[EventBatch], subset: Optional[Sequence[int]] = None) -> EventBatch:
    """Re-group events from batches (evt column renumbered to the new group)."""
    pieces = []
    for b in batches:
        idx = range(b.n_events) if subset is None else subset
        for e in idx:
            pieces.append((b, e))
    L = batches[0].n_layers
    xs, ys, zs, ls, ds, hs, es = [], [], [], [], [], [], []
    ev_off = np.zeros(len(pieces), np.uint32)
    l_n = np.zeros(len(pieces) * L, np.uint32)
    l_off = np.zeros(len(pieces) * L, np.uint32)
    total = 0
    for k, (b, e) in enumerate(pieces):
        s = int(b.event_offset[e])
        n = int(b.layer_nhits[e * L:(e + 1) * L].sum())
        sl = slice(s, s + n)
        xs.append(b.x[sl]); ys.append(b.y[sl]); zs.append(b.z[sl])
        ls.append(b.layer[sl]); ds.append(b.det_id[sl]); hs.append(b.hit_id[sl])
        es.append(np.full(n, k, np.uint32))
        ev_off[k] = total
        l_n[k * L:(k + 1) * L] = b.layer_nhits[e * L:(e + 1) * L]
        l_off[k * L:(k + 1) * L] = b.layer_offset[e * L:(e + 1) * L]
        total += n

    def cat(a, dt):
        return np.ascontiguousarray(np.concatenate(a).astype(dt)) if a else np.zeros(0, dt)

    return EventBatch(cat(xs, np.float32), cat(ys, np.float32), cat(zs, np.float32),
                      cat(ls, np.uint32), cat(ds, np.uint32), cat(hs, np.uint32),
                      cat(es, np.uint32), ev_off, l_n, l_off, L)
