"""Physics evaluation of the seeding output against simulated truth:
efficiency, fake rate and clone rate per (event, layer triplet), merged over
events, with eta and pT histograms, restating
src/triplet_builder/PhysicsRecord.{h,cpp} (fillData :81-218, merge :232-251,
csvDump :253-370, circle fit :41-79, TIP :287-304).

Classification of one (event, layer triplet)'s triplets (fillData):
* valid: the three hits carry one simulated-track id; a findable track's
  first valid triplet counts as found ("valid"), every further one as a
  clone; a valid triplet of a non-findable track is "misc";
* fake: the hit ids differ;
* missed: findable tracks with no valid triplet, binned by the eta of their
  innermost and outermost hit and the pT of the circle through the innermost,
  middle and outermost hit.
efficiency = found / findable, fake rate = fakes / triplets, clone rate =
clones / triplets (a zero denominator counts as 1, Utils::clamp).

Records of one layer triplet merge with the parallel mean/variance update the
reference uses (each record weighs n = 1, so the merged values are the mean
over events and its standard deviation).  Histogram bins merge the same way;
a bin that a merge creates starts from a default bin of weight n = 1, so its
fake and clone means are scaled by n / (n + 1) as in the reference
(tBinnedData::operator+=, :10-29, on a std::map default entry).

Arithmetic follows the reference's types: hit coordinates are float32 (and
x*x + y*y of the circle fit is a float32 expression), the fit and the
kinematics are double, means and variances long double (numpy.longdouble,
the same 80-bit type on x86-64 Linux); CSV values print with six significant
digits, as a default-formatted C++ stream does.

Triplet hit indexes address the grid-sorted hits, which the reference copies
back to the host after the grid build (GridBuilder.cpp:151): pass the sorted
columns (the chain's fetch returns them).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Mapping, Optional, Sequence, Tuple

import numpy as np

LD = np.longdouble

# PhysicsRecord.h:138-145
BZ = 3.8112  # T
Q = 1.602177e-19  # C
C_LIGHT = 2.998e8  # m/s
GEV_C = 5.344286e-19  # J s / m per GeV/c
BIN_WIDTH = 0.1


def _clamp(n: int) -> int:
    return n if n > 0 else 1


def _fmt(v) -> str:
    return "%g" % float(v)


def circle_params(p1, p2, p3) -> Tuple[float, float, float]:
    """(center x, center y, radius) of the circle through three (x, y) points
    by the paraboloid-lifting fit of Strandlie et al. (PhysicsRecord.cpp:41-79)."""
    with np.errstate(all="ignore"):
        pts = []
        for p in (p1, p2, p3):
            x, y = np.float32(p[0]), np.float32(p[1])
            pts.append((np.float64(x), np.float64(y), np.float64(np.float32(x * x + y * y))))
        (x1, y1, w1), (x2, y2, w2), (x3, y3, w3) = pts
        ax, ay, az = x2 - x1, y2 - y1, w2 - w1
        bx, by, bz = x3 - x1, y3 - y1, w3 - w1
        nx, ny, nz = ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx
        value = np.sqrt(nx * nx + ny * ny + nz * nz)
        nx, ny, nz = nx / value, ny / value, nz / value
        cx, cy = (-nx) / (2 * nz), (-ny) / (2 * nz)
        c = -(nx * x1 + ny * y1 + nz * w1)
        r = np.sqrt((1 - nz * nz - 4 * c * nz) / (4 * nz * nz))
    return float(cx), float(cy), float(r)


def pt_of_radius(radius_cm: float) -> float:
    """pT [GeV/c] of a helix of the given transverse radius [cm] (:175-177)."""
    return Q * BZ * (radius_cm * 1e-2) / GEV_C


def transverse_pt(p1, p2, p3) -> float:
    return pt_of_radius(circle_params(p1, p2, p3)[2])


def eta(inner, outer) -> float:
    """Pseudorapidity of the line inner -> outer (:160-168)."""
    with np.errstate(all="ignore"):
        px = np.float64(np.float32(outer[0]) - np.float32(inner[0]))
        py = np.float64(np.float32(outer[1]) - np.float32(inner[1]))
        pz = np.float64(np.float32(outer[2]) - np.float32(inner[2]))
        return float(np.arcsinh(pz / np.sqrt(px * px + py * py)))


def tip(p1, p2, p3) -> float:
    """Transverse impact parameter: distance from the origin to the circle's
    point of closest approach (:287-304)."""
    cx, cy, r = circle_params(p1, p2, p3)
    with np.errstate(all="ignore"):
        vx, vy = np.float64(-cx), np.float64(-cy)
        value = np.sqrt(vx * vx + vy * vy)
        vx, vy = vx / value, vy / value
        ax, ay = cx + r * vx, cy + r * vy
        return float(np.sqrt(ax * ax + ay * ay))


def bin_of(v: float, width: float = BIN_WIDTH) -> float:
    with np.errstate(all="ignore"):
        return float(np.float64(width) * np.floor(np.float64(v) / np.float64(width)))


@dataclass
class BinnedData:
    """tBinnedData (PhysicsRecord.h:47-72)."""
    valid: int = 0
    clones: int = 0
    fake: int = 0
    missed: int = 0
    misc: int = 0
    efficiency_mean: LD = LD(0)
    efficiency_var: LD = LD(0)
    fake_rate_mean: LD = LD(0)
    fake_rate_var: LD = LD(0)
    clone_rate_mean: LD = LD(0)
    clone_rate_var: LD = LD(0)
    n: int = 1

    def fill(self) -> None:
        tot = self.valid + self.fake + self.clones + self.misc
        self.efficiency_mean = LD(self.valid / _clamp(self.valid + self.missed))
        self.fake_rate_mean = LD(self.fake / _clamp(tot))
        self.clone_rate_mean = LD(self.clones / _clamp(tot))
        self.n = tot

    def to_var(self, x) -> LD:
        return np.sqrt(LD(x) / LD(_clamp(self.n - 1))) if self.n > 1 else LD(0)

    def __iadd__(self, c: "BinnedData") -> "BinnedData":
        self.valid += c.valid
        self.clones += c.clones
        self.fake += c.fake
        self.missed += c.missed
        self.n += c.n
        delta = c.efficiency_mean - self.efficiency_mean
        w = LD(c.valid + c.missed)
        self.efficiency_mean += w * delta / LD(_clamp(self.valid + self.missed))
        self.efficiency_var += w * delta * (c.efficiency_mean - self.efficiency_mean) + c.efficiency_var
        delta = c.fake_rate_mean - self.fake_rate_mean
        self.fake_rate_mean += LD(c.n) * delta / LD(_clamp(self.n))
        self.fake_rate_var += LD(c.n) * delta * (c.fake_rate_mean - self.fake_rate_mean) + c.fake_rate_var
        delta = c.clone_rate_mean - self.clone_rate_mean
        self.clone_rate_mean += LD(c.n) * delta / LD(_clamp(self.n))
        self.clone_rate_var += LD(c.n) * delta * (c.clone_rate_mean - self.clone_rate_mean) + c.clone_rate_var
        return self

    def csv(self) -> str:
        return " ".join([str(self.valid), str(self.fake), str(self.clones), str(self.missed), str(self.misc)]
                        + [_fmt(v) for v in (self.efficiency_mean, self.to_var(self.efficiency_var),
                                             self.fake_rate_mean, self.to_var(self.fake_rate_var),
                                             self.clone_rate_mean, self.to_var(self.clone_rate_var))])


class Histogram(dict):
    """tHistogram: bin lower edge -> BinnedData, ordered by bin on output."""

    def __init__(self, bin_width: float = BIN_WIDTH):
        super().__init__()
        self.bin_width = bin_width

    def at(self, b: float) -> BinnedData:
        if b not in self:
            self[b] = BinnedData()
        return self[b]

    def fill(self) -> None:
        for v in self.values():
            v.fill()

    def merge(self, other: "Histogram") -> None:
        """mergeHistograms (:220-230)."""
        for b, d in sorted(other.items()):
            if self.bin_width == other.bin_width:
                self.at(b).__iadd__(d)
            else:
                centre = b * other.bin_width + other.bin_width / 2
                self.at(self.bin_width * math.floor(centre / self.bin_width)).__iadd__(d)

    def csv(self) -> str:
        head = ("bin valid fake clones missed misc efficiencyMean efficiencyVar fakeRateMean "
                "fakeRateVar cloneRateMean cloneRateVar\n")
        return head + "".join(f"{_fmt(b)} {self[b].csv()}\n" for b in sorted(self))


@dataclass
class SortedHits:
    """The grid-sorted hit columns the triplets index."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    hit_id: np.ndarray

    def point(self, i: int):
        return (self.x[i], self.y[i], self.z[i])


TruthTrack = np.ndarray  # [n, 3] float32 (x, y, z) of a findable track's hits, innermost layer first


@dataclass
class PhysicsRecord:
    event: int
    layer_triplet: int
    efficiency_mean: LD = LD(0)
    efficiency_var: LD = LD(0)
    fake_rate_mean: LD = LD(0)
    fake_rate_var: LD = LD(0)
    clone_rate_mean: LD = LD(0)
    clone_rate_var: LD = LD(0)
    n: int = 1
    pt: Histogram = field(default_factory=Histogram)
    eta: Histogram = field(default_factory=Histogram)
    counts: Dict[str, int] = field(default_factory=dict)  # this record's own classification totals

    def to_var(self, x) -> LD:
        return np.sqrt(LD(x) / LD(_clamp(self.n - 1))) if self.n > 1 else LD(0)

    def fill_data(self, tracklets: np.ndarray, tracklet_offsets: np.ndarray, mc_truth: Mapping[int, TruthTrack],
                  hits: SortedHits, n_layer_triplets: int) -> None:
        """PhysicsRecord::fillData (:81-218)."""
        k = self.event * n_layer_triplets + self.layer_triplet
        lo, hi = int(tracklet_offsets[k]), int(tracklet_offsets[k + 1])
        found: set = set()
        fakes = clones = valid = misc = 0
        hid = hits.hit_id
        for i in range(lo, hi):
            a, b, c = (int(v) for v in tracklets[i])
            p1, p2, p3 = hits.point(a), hits.point(b), hits.point(c)
            e_bin, p_bin = bin_of(eta(p1, p3)), bin_of(transverse_pt(p1, p2, p3))
            if hid[a] == hid[b] and hid[a] == hid[c]:
                t = int(hid[a])
                if t in mc_truth:
                    if t not in found:
                        found.add(t)
                        valid += 1
                        self.eta.at(e_bin).valid += 1
                        self.pt.at(p_bin).valid += 1
                    else:
                        clones += 1
                        self.eta.at(e_bin).clones += 1
                        self.pt.at(p_bin).clones += 1
                else:
                    misc += 1
                    self.eta.at(e_bin).misc += 1
                    self.pt.at(p_bin).misc += 1
            else:
                fakes += 1
                self.eta.at(e_bin).fake += 1
                self.pt.at(p_bin).fake += 1
        missed = 0
        for t in sorted(mc_truth):
            if t in found:
                continue
            trk = mc_truth[t]
            inner, middle, outer = trk[0], trk[len(trk) // 2], trk[len(trk) - 1]
            self.eta.at(bin_of(eta(inner, outer))).missed += 1
            self.pt.at(bin_of(transverse_pt(inner, middle, outer))).missed += 1
            missed += 1
        n_found = hi - lo
        self.efficiency_mean = LD(len(found) / _clamp(len(mc_truth)))
        self.fake_rate_mean = LD(fakes / _clamp(n_found))
        self.clone_rate_mean = LD(clones / _clamp(n_found))
        self.eta.fill()
        self.pt.fill()
        self.counts = dict(triplets=n_found, valid=valid, clones=clones, fake=fakes, misc=misc, missed=missed,
                           findable=len(mc_truth))

    def merge(self, c: "PhysicsRecord") -> None:
        """PhysicsRecord::merge (:232-251)."""
        self.eta.merge(c.eta)
        self.pt.merge(c.pt)
        self.n += c.n
        for m, v in (("efficiency_mean", "efficiency_var"), ("fake_rate_mean", "fake_rate_var"),
                     ("clone_rate_mean", "clone_rate_var")):
            delta = getattr(c, m) - getattr(self, m)
            setattr(self, m, getattr(self, m) + LD(c.n) * delta / LD(_clamp(self.n)))
            setattr(self, v, getattr(self, v) + LD(c.n) * delta * (getattr(c, m) - getattr(self, m)) + getattr(c, v))
        for key, val in c.counts.items():
            self.counts[key] = self.counts.get(key, 0) + val

    def copy(self) -> "PhysicsRecord":
        r = PhysicsRecord(self.event, self.layer_triplet, self.efficiency_mean, self.efficiency_var,
                          self.fake_rate_mean, self.fake_rate_var, self.clone_rate_mean, self.clone_rate_var,
                          self.n, Histogram(self.pt.bin_width), Histogram(self.eta.bin_width), dict(self.counts))
        for src, dst in ((self.pt, r.pt), (self.eta, r.eta)):
            for b, d in src.items():
                dst[b] = BinnedData(**d.__dict__)
        return r

    def csv(self, output_dir: Optional[str] = None) -> str:
        """One CSV row (layerTriplet n eff effVar fr frVar cr crVar); with an
        output directory also eta_<lt>.csv and pt_<lt>.csv (:253-285)."""
        row = " ".join([str(self.layer_triplet), str(self.n),
                        _fmt(self.efficiency_mean), _fmt(self.to_var(self.efficiency_var)),
                        _fmt(self.fake_rate_mean), _fmt(self.to_var(self.fake_rate_var)),
                        _fmt(self.clone_rate_mean), _fmt(self.to_var(self.clone_rate_var))])
        if output_dir:
            with open(os.path.join(output_dir, f"eta_{self.layer_triplet}.csv"), "w") as f:
                f.write(self.eta.csv())
            with open(os.path.join(output_dir, f"pt_{self.layer_triplet}.csv"), "w") as f:
                f.write(self.pt.csv())
        return row


class PhysicsRecords:
    """Records merged per layer triplet in order of first appearance
    (PhysicsRecords::addRecord, :306-318)."""

    def __init__(self):
        self.records: List[PhysicsRecord] = []

    def add_record(self, r: PhysicsRecord) -> None:
        for mine in self.records:
            if mine.layer_triplet == r.layer_triplet:
                mine.merge(r)
                return
        self.records.append(r.copy())

    def merge(self, other: "PhysicsRecords") -> None:
        for r in other.records:
            self.add_record(r)

    def total(self) -> PhysicsRecord:
        """All layer triplets merged, labelled len(records) + 1 (:330-340)."""
        if not self.records:
            raise ValueError("no physics records")
        t = self.records[0].copy()
        t.layer_triplet = len(self.records) + 1
        for r in self.records[1:]:
            t.merge(r)
        return t

    def csv(self, output_dir: Optional[str] = None) -> str:
        """physics.<config>.csv: header, one row per layer triplet, the total."""
        total = self.total()
        out = ["layerTriplet n eff effVar fr frVar cr crVar"]
        out += [r.csv(output_dir) for r in self.records]
        out.append(total.csv(output_dir))
        return "\n".join(out) + "\n"


def mc_truth(event, track_hits: Mapping[int, Sequence[int]]) -> Dict[int, TruthTrack]:
    """A findable-track list (track id -> hit indexes of the event file, in
    layer order) as coordinates."""
    return {t: np.stack([event.x[list(ix)], event.y[list(ix)], event.z[list(ix)]], axis=1)
            for t, ix in track_hits.items()}


def evaluate(chain: Mapping[str, np.ndarray], events: Sequence, tracks: Sequence[Mapping[int, Sequence[int]]],
             n_layer_triplets: int, records: Optional[PhysicsRecords] = None) -> PhysicsRecords:
    """Physics records of one event group (TripletBuilder.cpp:272-279):
    `chain` holds the sorted hit columns x, y, z, hit_id, the triplets and
    their offsets; events and tracks are the group's events and findable
    tracks (cmsdata.load_events)."""
    hits = SortedHits(chain["x"], chain["y"], chain["z"], chain["hit_id"])
    records = records if records is not None else PhysicsRecords()
    for e, (ev, tr) in enumerate(zip(events, tracks)):
        truth = mc_truth(ev, tr)
        for p in range(n_layer_triplets):
            r = PhysicsRecord(e, p)
            r.fill_data(chain["tracklets"], chain["tracklet_offsets"], truth, hits, n_layer_triplets)
            records.add_record(r)
    return records
