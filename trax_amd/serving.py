"""Continuous-batching seeding server.

Events arrive as a stream (e.g. Poisson).  The server owns a fixed set of
slots; each slot is a HIP stream with its own trax context and resident
device buffers (hits, sorted hits, grid, pairs, candidates, triplets) sized
so that all slots together stay inside the device-memory budget.  Whenever a
slot is free, the events waiting at that moment are admitted into it as one
dynamic batch (1..max_batch events, limited by the slot's hit capacity), and
the whole chain runs on that slot's stream while other slots are still busy,
so batches overlap on the device.  Latency of an event = completion of its
batch - its arrival; results are split back per event.

An event's triplets do not depend on the batch it is served in: the chain
treats every (event, layer) segment independently.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import events as ev
from . import runtime as rt


HIT_ROWS = ("x", "y", "z", "layer", "det_id", "hit_id")  # rows of a request's packed hit block


@dataclass
class Request:
    rid: int
    arrival: float  # seconds, relative to the start of the run
    batch: ev.EventBatch  # one event
    done: float = -1.0
    triplets: Optional[np.ndarray] = None  # (n, 3) int32 input-hit indices within the event
    hits: Optional[np.ndarray] = None  # [6, n] uint32: HIT_ROWS (x, y, z as float32 bits)
    layout: Optional[np.ndarray] = None  # [2, L] uint32: per-layer hit counts and offsets

    def __post_init__(self):
        # the request's wire form: its hit columns as one contiguous block
        if self.hits is None:
            b = self.batch
            self.hits = np.stack([b.x.view(np.uint32), b.y.view(np.uint32), b.z.view(np.uint32),
                                  b.layer, b.det_id, b.hit_id]).astype(np.uint32, copy=False)
        if self.layout is None:
            self.layout = np.stack([self.batch.layer_nhits, self.batch.layer_offset]).astype(np.uint32)


def pack_requests(reqs: Sequence["Request"], hits: np.ndarray, evt: np.ndarray, layout: np.ndarray,
                  event_offset: np.ndarray) -> Tuple[int, np.ndarray]:
    """Lay a batch of requests out as the chain reads it (event-major hit
    columns, the event column renumbered 0..E-1, per-(event, layer) counts
    and event-relative offsets, per-event offsets) in the given staging
    blocks: hits [6, >= n] uint32, evt [>= n], layout [2, >= E * L],
    event_offset [>= E].  Returns (n_hits, per-request first hit)."""
    E = len(reqs)
    sizes = np.fromiter((r.hits.shape[1] for r in reqs), np.int64, E)
    offsets = np.zeros(E, np.int64)
    np.cumsum(sizes[:-1], out=offsets[1:])
    n = int(sizes.sum())
    np.concatenate([r.hits for r in reqs], axis=1, out=hits[:, :n])
    evt[:n] = np.repeat(np.arange(E, dtype=evt.dtype), sizes)
    m = E * reqs[0].layout.shape[1] if E else 0
    np.concatenate([r.layout for r in reqs], axis=1, out=layout[:, :m])
    event_offset[:E] = offsets
    return n, offsets


class Slot:
    max_regrow = 4  # buffer regrowths per batch before a slot gives up

    def __init__(self, sc: ev.SeedConfig, max_events: int, max_hits: int, pair_cap: int,
                 cand_cap: int, trk_cap: int, device="cuda"):
        import torch
        self.torch = torch
        self.device = device
        self.regrowths = 0
        self.stream = torch.cuda.Stream(device=device)
        self.ctx = rt.Context(torch.device(device).index or 0)
        # only the triplets are served: the chain runs the fused prediction +
        # filter and keeps no candidate records
        self.bufs = rt.ChainBuffers(sc, max_events, max_hits, pair_cap, cand_cap, trk_cap, device,
                                    candidates=False)
        n = max(max_hits, 1)
        L = max(1, sc.n_layers)
        # hit rows, the event column and the layout, each one block on the
        # device and one pinned block on the host
        self.dev_hits = torch.empty((len(HIT_ROWS), n), dtype=torch.int32, device=device)
        self.dev_evt = torch.empty(n, dtype=torch.int32, device=device)
        self.dev_lay = torch.empty((2, max_events * L), dtype=torch.int32, device=device)
        self.dev_eoff = torch.empty(max_events, dtype=torch.int32, device=device)
        self.host_hits = torch.empty((len(HIT_ROWS), n), dtype=torch.int32).pin_memory()
        self.host_evt = torch.empty(n, dtype=torch.int32).pin_memory()
        self.host_lay = torch.empty((2, max_events * L), dtype=torch.int32).pin_memory()
        self.host_eoff = torch.empty(max_events, dtype=torch.int32).pin_memory()
        self.dev = {k: (self.dev_hits[i].view(torch.float32) if k in "xyz" else self.dev_hits[i])
                    for i, k in enumerate(HIT_ROWS)}
        self.dev["evt"] = self.dev_evt
        # results: triplets as input-hit indices (gathered on the device) and
        # the per-(event, layer triplet) offsets, copied into pinned blocks
        self.host_trk = torch.empty((0, 3), dtype=torch.int32).pin_memory()
        self.host_toff = torch.empty(max_events * max(1, len(sc.triplets)) + 1, dtype=torch.int32).pin_memory()
        self.requests: List[Request] = []
        self.done_event = None
        self.db: Optional[rt.DeviceBatch] = None
        self.offsets = np.zeros(0, np.int64)  # first hit of each request in the batch
        self.n_hits = 0

    @property
    def busy(self) -> bool:
        return bool(self.requests)

    def nbytes(self) -> int:
        b = self.bufs
        tot = sum(t.numel() * t.element_size() for t in b.sorted.values())
        tot += b.grid.numel() * 4 + b.ws_mem.numel() + b.pairs.numel() * 4
        tot += b.cands.numel() * 4 if b.cands is not None else 0
        tot += sum(t.numel() * 4 for t in b.t)
        tot += 4 * (self.dev_hits.numel() + self.dev_evt.numel() + self.dev_lay.numel() + self.dev_eoff.numel())
        return int(tot)

    def launch(self, reqs: List[Request]) -> None:
        """Gather the requests' packed hit blocks straight into this slot's
        pinned staging blocks (one vectorised concatenation per block, event
        index renumbered), upload them and enqueue the chain on the slot's
        stream."""
        torch = self.torch
        L = self.bufs.sc.n_layers
        E = len(reqs)
        n, offsets = pack_requests(reqs, self.host_hits.numpy().view(np.uint32), self.host_evt.numpy(),
                                   self.host_lay.numpy().view(np.uint32), self.host_eoff.numpy())
        with torch.cuda.stream(self.stream):
            for i in range(len(HIT_ROWS)):
                self.dev_hits[i, :n].copy_(self.host_hits[i, :n], non_blocking=True)
            self.dev_evt[:n].copy_(self.host_evt[:n], non_blocking=True)
            self.dev_lay[:, :E * L].copy_(self.host_lay[:, :E * L], non_blocking=True)
            self.dev_eoff[:E].copy_(self.host_eoff[:E], non_blocking=True)
            cols = dict(self.dev)
            cols.update(event_offset=self.dev_eoff, layer_nhits=self.dev_lay[0], layer_offset=self.dev_lay[1])
            self.db = rt.DeviceBatch(E, n, cols)
            rt.seed_chain(self.ctx, self.bufs, self.db, stream=self.stream)
            self.done_event = torch.cuda.Event()
            self.done_event.record(self.stream)
        self.requests = reqs
        self.offsets = offsets
        self.n_hits = n

    def ready(self) -> bool:
        return self.busy and self.done_event.query()

    def finish(self, now: float) -> List[Request]:
        """Collect the results of the finished batch (its stream is idle)."""
        reqs = self.requests
        flags = self.ctx.flags(stream=self.stream)
        if flags & rt.FLAG_LAYOUT:
            raise rt.TraxError("inconsistent event layout")
        for _ in range(self.max_regrow):
            if not flags:
                break
            # outputs overflowed: grow this slot's buffers and rerun
            # synchronously.  After a pair overflow the candidate and triplet
            # totals came from truncated pairs, so the rerun may overflow a
            # later stage in turn: hence more than one round.
            n_p, n_c, n_t = rt.totals(self.bufs)
            b = self.bufs
            self.bufs = rt.ChainBuffers(b.sc, b.max_events, b.max_hits, max(n_p, b.pair_capacity),
                                        max(n_c, b.cand_capacity), max(n_t, b.tracklet_capacity),
                                        device=self.device, candidates=b.candidates)
            self.regrowths += 1
            with self.torch.cuda.stream(self.stream):
                rt.seed_chain(self.ctx, self.bufs, self.db, stream=self.stream)
            self.stream.synchronize()
            flags = self.ctx.flags(stream=self.stream)
            if flags & rt.FLAG_LAYOUT:
                raise rt.TraxError("inconsistent event layout")
        if flags:
            raise rt.TraxError("slot still overflows after regrowing")
        torch = self.torch
        P = len(self.bufs.sc.triplets)
        E = len(reqs)
        n_t = int(rt.totals(self.bufs)[2])
        if self.host_trk.shape[0] < n_t:
            self.host_trk = torch.empty((n_t + n_t // 4 + 1024, 3), dtype=torch.int32).pin_memory()
        with torch.cuda.stream(self.stream):
            if n_t:
                # input-hit index of every triplet hit, gathered on the device:
                # only the results cross PCIe
                idx = torch.stack([t[:n_t] for t in self.bufs.t], dim=1).long()
                self.host_trk[:n_t].copy_(self.bufs.sorted["orig_index"][idx], non_blocking=True)
            self.host_toff[:E * P + 1].copy_(self.bufs.tracklet_offsets[:E * P + 1], non_blocking=True)
        self.stream.synchronize()
        offs = self.host_toff[:E * P + 1].numpy()
        trk = self.host_trk[:n_t].numpy()
        for e, r in enumerate(reqs):
            # a new int32 array per request (the pinned block is reused)
            r.triplets = trk[offs[e * P]:offs[(e + 1) * P]] - int(self.offsets[e])
            r.done = now
        self.requests = []
        return reqs


@dataclass
class ServeStats:
    n_events: int
    wall_s: float
    events_per_sec: float
    p50_ms: float
    p99_ms: float
    batches: int
    mean_batch: float
    max_batch: int
    slots: int
    budget_bytes: int
    used_bytes: int
    batch_sizes: List[int] = field(default_factory=list)


class SeedServer:
    def __init__(self, sc: ev.SeedConfig, budget_bytes: int, max_batch: int = 512, n_slots: int = 3,
                 slot_hits: int = 1_000_000, pairs_per_hit: int = 48, device="cuda",
                 triplets_per_pair: float = 0.25):
        self.sc = sc
        self.max_batch = max_batch
        self.budget = int(budget_bytes)
        self.slots: List[Slot] = []
        pair_cap = pairs_per_hit * slot_hits
        trk_cap = max(1, int(pair_cap * triplets_per_pair))
        for _ in range(n_slots):
            s = Slot(sc, max_batch, slot_hits, pair_cap, 2 * pair_cap, trk_cap, device)
            if sum(x.nbytes() for x in self.slots) + s.nbytes() > self.budget:
                del s
                break
            self.slots.append(s)
        if not self.slots:
            raise rt.TraxError("device-memory budget too small for one slot")
        self.slot_hits = slot_hits

    def used_bytes(self) -> int:
        return sum(s.nbytes() for s in self.slots)

    def serve(self, requests: Sequence[Request], realtime: bool = True) -> ServeStats:
        """Serve requests in arrival order; realtime=False admits them all at t=0."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: List[Request] = []
        finished: List[Request] = []
        sizes: List[int] = []
        t0 = time.perf_counter()
        i = 0
        while len(finished) < len(pending):
            now = time.perf_counter() - t0
            while i < len(pending) and (not realtime or pending[i].arrival <= now):
                waiting.append(pending[i])
                i += 1
            progressed = False
            for s in self.slots:
                if s.busy and s.ready():
                    finished.extend(s.finish(time.perf_counter() - t0))
                    progressed = True
            for s in self.slots:
                if not s.busy and waiting:
                    take, hits = [], 0
                    while waiting and len(take) < self.max_batch:
                        h = waiting[0].batch.n_hits
                        if hits + h > self.slot_hits:
                            break
                        take.append(waiting.pop(0))
                        hits += h
                    if not take:
                        raise rt.TraxError("an event larger than a slot")
                    s.launch(take)
                    sizes.append(len(take))
                    progressed = True
            if not progressed:
                time.sleep(0.0002)
        wall = max(r.done for r in finished) - min(r.arrival for r in finished)
        lat = np.array([(r.done - r.arrival) * 1e3 for r in finished])
        return ServeStats(len(finished), wall, len(finished) / wall, float(np.percentile(lat, 50)),
                          float(np.percentile(lat, 99)), len(sizes), float(np.mean(sizes)),
                          int(max(sizes)), len(self.slots), self.budget, self.used_bytes(), sizes)


def poisson_requests(w: ev.Workload, n: int, rate_hz: float, seed: int = ev.SEED_BASE) -> List[Request]:
    """n events of workload w arriving as a Poisson process of the given rate."""
    rng = np.random.default_rng([seed, 0xA77, w.config_id])
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    return [Request(k, float(t[k]), ev.make_batch(w, events=[k], seed=seed)) for k in range(n)]


def _check_against_plain(ctx, sc: ev.SeedConfig, reqs: Sequence[Request]) -> int:
    """Every request's served triplets equal the plain method's (one event,
    stage by stage with host reads) on the same event; returns the count."""
    for r in reqs:
        bufs, db = rt.run_batch(ctx, sc, r.batch)
        rt.run_plain_event(ctx, bufs, db)
        alone = rt.fetch_triplets(bufs, 1, r.batch.n_hits)
        t = alone["orig_index"][alone["tracklets"]].astype(np.int64)
        key = lambda a: a[np.lexsort((a[:, 2], a[:, 1], a[:, 0]))] if len(a) else a.reshape(0, 3)
        if not np.array_equal(key(r.triplets.reshape(-1, 3)), key(t.reshape(-1, 3))):
            raise rt.TraxError(f"request {r.rid}: served triplets differ from the plain path")
    return len(reqs)


def main(argv=None) -> int:
    """Serving benchmark: a Poisson stream of the PU20/PU200 mix at several
    arrival rates (and all requests at t = 0), one line of JSON per rate."""
    import argparse
    import json

    ap = argparse.ArgumentParser(prog="python -m trax_amd.serving")
    ap.add_argument("--events", type=int, default=4096)
    ap.add_argument("--rates", default="5000,20000,50000,0",
                    help="arrival rates in events/s; 0 = every request at t = 0")
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--max-batch", type=int, default=512)
    ap.add_argument("--slots", type=int, default=3)
    ap.add_argument("--slot-hits", type=int, default=2_000_000)
    ap.add_argument("--check", type=int, default=64, help="served events checked against the plain path")
    args = ap.parse_args(argv)
    import torch
    w = ev.config_workload(4, args.events)
    srv = SeedServer(w.seed_config, budget_bytes=int(args.budget_gb * (1 << 30)), max_batch=args.max_batch,
                     n_slots=args.slots, slot_hits=args.slot_hits)
    ctx = rt.Context(torch.cuda.current_device())
    for rate in [float(x) for x in args.rates.split(",")]:
        srv.serve(poisson_requests(w, 64, 1e9, seed=ev.SEED_BASE + 1), realtime=False)  # warm-up
        reqs = poisson_requests(w, args.events, rate if rate > 0 else 1e9)
        st = srv.serve(reqs, realtime=rate > 0)
        checked = _check_against_plain(ctx, w.seed_config, reqs[:: max(1, len(reqs) // args.check)])
        print(json.dumps({"workload": w.name, "arrival_rate_hz": rate or None, "events": st.n_events,
                          "events_per_sec": st.events_per_sec, "p50_ms": st.p50_ms, "p99_ms": st.p99_ms,
                          "batches": st.batches, "mean_batch": st.mean_batch, "max_batch": st.max_batch,
                          "slots": st.slots, "budget_gb": st.budget_bytes / (1 << 30),
                          "used_gb": st.used_bytes / (1 << 30), "checked_vs_plain": checked}), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
