"""Autotuner for the seeding chain.

Searches, per workload, the grid binning (phi x z sectors) and the launch
shape of every stage (waves per grid-build block, threads per block of the
two-phase pair and prediction kernels, scan tile size, and the fused
prediction + filter walk's lane buffer depth, waves per block and waves per
SIMD), timing each stage with HIP events on the launch stream.  Stages are
tuned independently for a given binning (their launch shapes do not change
their inputs); binnings are compared on the whole chain (the fused chain the
benchmark and the server run) with each stage at its best shape.  Results
are written as JSON and can be loaded into a context with `apply`.

The binning changes the grid and so, in edge cases, the pair and triplet
sets (the reference's window walk depends on the cells); the parity tests
run the oracle at the tuned binning.
"""
from __future__ import annotations

import json
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from . import events as ev
from . import runtime as rt

WAVES = (0, 1, 2, 4, 8)  # 0: chosen from the mean segment size
BLOCKS = (64, 128, 256, 512)
SCAN_TILES = (4096, 8192, 16384)
FUSED = rt.FUSED_SHAPES  # (lane buffer, waves per block, waves per SIMD)
PHI_BINS = (64, 128, 256)
Z_BINS = (64, 128, 256, 512)
# the binning space per workload: the uneven pile-up-20..200 batch is tuned
# over 64-256 phi x 32-128 z sectors
SPACES = {2: ((64, 128, 256), (32, 64, 128))}


@dataclass
class TuneResult:
    workload: str
    n_phi: int
    n_z: int
    launch: Dict[str, int]
    chain_ms: float
    stage_ms: Dict[str, float]
    table: List[dict] = field(default_factory=list)

    def to_json(self) -> str:
        return json.dumps(self.__dict__, indent=1)


def _time(fn, reps, torch) -> float:
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune_binning(ctx: rt.Context, sc: ev.SeedConfig, batch: ev.EventBatch, reps: int = 3):
    """Best launch shape per stage at this binning; returns (launch, stage_ms, chain_ms)."""
    import torch
    base = dict(grid_waves=0, pairs_block=256, predict_block=256, scan_tile=4096,
                fused_buffer=0, fused_block_waves=0, fused_waves_per_eu=0)
    ctx.set_launch(**base)
    bufs, db = rt.run_batch(ctx, sc, batch)
    fbufs, fdb = rt.run_batch(ctx, sc, batch, candidates=False)
    E = db.n_events
    best = dict(base)
    stage_ms: Dict[str, float] = {}

    def sweep(key, values, fn, name):
        times = {}
        for v in values:
            cfg = dict(best)
            if isinstance(key, tuple):
                cfg.update(zip(key, v))
            else:
                cfg[key] = v
            ctx.set_launch(**cfg)
            times[v] = _time(fn, reps, torch)
        v = min(times, key=times.get)
        if isinstance(key, tuple):
            best.update(zip(key, v))
        else:
            best[key] = v
        stage_ms[name] = times[v]
        return times

    sweep("grid_waves", WAVES, lambda: rt.grid_build(ctx, bufs, db), "grid_build")
    sweep("scan_tile", SCAN_TILES, lambda: rt.pairs_count(ctx, bufs, E), "pairs_count_scan")
    sweep("pairs_block", BLOCKS, lambda: (rt.pairs_count(ctx, bufs, E), rt.pairs_store(ctx, bufs, E)),
          "pairs_two_phase")
    sweep("predict_block", BLOCKS,
          lambda: (rt.predict_count(ctx, bufs, E), rt.predict_store(ctx, bufs, E)), "predict_two_phase")
    sweep(("fused_buffer", "fused_block_waves", "fused_waves_per_eu"), FUSED,
          lambda: rt.triplet_predict_filter(ctx, fbufs, E), "predict_filter")
    stage_ms["pairs"] = _time(lambda: rt.pair_stage(ctx, bufs, E), reps, torch)
    stage_ms["predict"] = _time(lambda: rt.triplet_predict(ctx, bufs, E), reps, torch)
    ctx.set_launch(**best)
    stage_ms["filter"] = _time(lambda: rt.triplet_filter(ctx, bufs, E), reps, torch)
    chain_ms = _time(lambda: rt.seed_chain(ctx, fbufs, fdb), reps, torch)
    if ctx.flags():
        raise rt.TraxError("overflow during tuning")
    return best, stage_ms, chain_ms


def tune(ctx: rt.Context, w: ev.Workload, batch: Optional[ev.EventBatch] = None,
         phi_bins: Sequence[int] = PHI_BINS, z_bins: Sequence[int] = Z_BINS,
         reps: int = 3, log=None) -> TuneResult:
    batch = batch if batch is not None else ev.make_batch(w)
    table = []
    best: Optional[Tuple] = None
    for n_phi in phi_bins:
        for n_z in z_bins:
            sc = ev.SeedConfig(n_z=n_z, n_phi=n_phi, min_radius=w.seed_config.min_radius,
                               max_radius=w.seed_config.max_radius, triplets=w.seed_config.triplets,
                               min_pt=w.seed_config.min_pt)
            t0 = time.perf_counter()
            launch, st, chain = tune_binning(ctx, sc, batch, reps)
            row = {"n_phi": n_phi, "n_z": n_z, "chain_ms": chain, "stage_ms": st, "launch": launch,
                   "tune_s": time.perf_counter() - t0}
            table.append(row)
            if log:
                log(f"phi {n_phi} z {n_z}: chain {chain:.3f} ms {launch}")
            if best is None or chain < best[0]:
                best = (chain, n_phi, n_z, launch, st)
    chain, n_phi, n_z, launch, st = best
    return TuneResult(w.name, n_phi, n_z, launch, chain, st, table)


def apply(ctx: rt.Context, w: ev.Workload, result: TuneResult) -> ev.Workload:
    """Set the tuned launch shape on ctx and the tuned binning on the workload."""
    ctx.set_launch(**result.launch)
    w.seed_config.n_phi, w.seed_config.n_z = result.n_phi, result.n_z
    return w


def main():
    import argparse

    ap = argparse.ArgumentParser(description="autotune binning and launch shapes for a workload")
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--events", type=int, default=None)
    ap.add_argument("--phi", type=int, nargs="*", default=None)
    ap.add_argument("--z", type=int, nargs="*", default=None)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    w = ev.config_workload(args.config, args.events)
    phi_bins, z_bins = SPACES.get(args.config, (PHI_BINS, Z_BINS))
    ctx = rt.Context(0)
    res = tune(ctx, w, phi_bins=args.phi or phi_bins, z_bins=args.z or z_bins, reps=args.reps,
               log=lambda m: print(m, flush=True))
    with open(args.out, "w") as f:
        f.write(res.to_json())
    print(f"best: {res.n_phi} phi x {res.n_z} z, chain {res.chain_ms:.3f} ms, launch {res.launch}",
          flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
