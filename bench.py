"""Benchmark of the track-seeding chain on one MI355X.

One step = the whole chain (grid build -> pairs -> triplet prediction ->
triplet filter, with the prefix sums between them) over one batch of
synthetic events already resident in HBM, replayed from a captured hipGraph.
The default workload is the high-pile-up batch: 64 events at pile-up 200 on
six barrel layers with four layer triplets.

Printed as one JSON line: batched events/s (value), triplets/s, the plain
one-event-per-launch path on the same events, the roofline of the dominant
kernel (HBM bytes it must move / its measured duration) and the CPU oracle
baseline on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, help="workload configuration 0..3")
    ap.add_argument("--events", type=int, default=None)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: no plain path, no CPU baseline, a 256 MiB "
                         "copy kernel for counter calibration")
    ap.add_argument("--tune", default=None,
                    help="autotuning JSON to apply (trax_amd.autotune); default: trax_amd/tuned/<workload>.json "
                         "when it exists")
    ap.add_argument("--no-tune", action="store_true", help="library default launch shapes and binning")
    ap.add_argument("--event-file", default=None,
                    help="PEventContainer file (e.g. the reference's hits.ttBar.pb) instead of synthetic events")
    ap.add_argument("--geometry", default=None,
                    help="directory holding detectorRadius.dat and radiusDictionary.dat (with --event-file)")
    ap.add_argument("--pair-generator", choices=["beamspot", "sector"], default="beamspot",
                    help="pair stage: PairGeneratorBeamspot (default) or PairGeneratorSector")
    ap.add_argument("--unfused", action="store_true",
                    help="time the chain with separate prediction and filter stages (candidate "
                         "records stored) instead of the fused prediction + filter pass")
    return ap.parse_args()


def time_events(fn, reps: int, torch):
    """Average device duration of fn() on the current stream (HIP events)."""
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


# Algorithmic bytes per launch of each chain kernel: what it must read and
# write at least once (inputs read once, outputs written once), from the
# batch's counts.  c: hits, pairs, cands (candidate slots), triplets, cells.
ALGO_BYTES = {
    # 7 input columns read (28 B), 9 sorted columns (36 B), the (z, phi) pair
    # (8 B) and the (x, y, z, phi) quadruple (16 B) written, one uint32 per cell
    "grid_build": lambda c: 88 * c["hits"] + 4 * c["cells"],
    # x, y, z, phi of every hit and the cells read, one uint2 per pair written
    # (the stage's count and store launches share these bytes)
    "pairs": lambda c: 16 * c["hits"] + 4 * c["cells"] + 8 * c["pairs"],
    # sector pairs, count + store passes: z and phi of every hit read twice,
    # its count written and read, the cells read, one uint2 per pair written
    "pairs_sector": lambda c: 24 * c["hits"] + 4 * c["cells"] + 8 * c["pairs"],
    # pair read + count written per pair, z and phi of every hit and the cells
    # read, one stash slot per candidate slot written
    "predict_walk": lambda c: 16 * c["hits"] + 4 * c["cells"] + 12 * c["pairs"] + 4 * c["cands"],
    # scanned count per pair and stash slot per candidate read, uint2 written
    "predict_place": lambda c: 4 * c["pairs"] + 12 * c["cands"],
    # candidate and pair records and x, y, z of every hit read, mask words
    # and their counts written
    "filter_mask": lambda c: 8 * c["cands"] + 8 * c["pairs"] + 12 * c["hits"] + 12 * c["cands"] // 64,
    # mask words + scanned counts read; per triplet a candidate and a pair
    # read and three hit indexes written
    "filter_store": lambda c: 12 * c["cands"] // 64 + 28 * c["triplets"],
    # fused prediction + filter walk: one pair record read per pair (the
    # per-pair slot budgets are written only for the few tiles read back),
    # z and phi of every hit and the cells read, x, y, z of each candidate's
    # third hit read, one 16-byte (pair, hit1, hit2, hit3) record per
    # accepted triplet written (the phi survivors it hands to the cut kernel
    # are intermediate records of the fused stage and not counted)
    "predict_filter": lambda c: 16 * c["hits"] + 4 * c["cells"] + 8 * c["pairs"] + 12 * c["cands"]
                                + 16 * c["triplets"],
    # per triplet its 16-byte record read and three hit indexes written
    "triplet_place": lambda c: 28 * c["triplets"],
}


# The grid stage's bytes as the reference's three passes move them (count:
# 12 B/hit + 4 B/cell; scan: 8 B/cell; store: 28 B/hit read + 28 B/hit
# written + 4 B/cell), plus the input-index column this build also writes
# (8 B/hit): 76 B/hit + 16 B/cell.  This build writes the grid once, so the
# figure over-counts its own traffic; ALGO_BYTES["grid_build"] is its own
# minimum and is reported beside it.
def grid_stage_bytes(c):
    return 76 * c["hits"] + 16 * c["cells"]


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when no summary covers this workload and kernel."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = dict(kv.split("=", 1) for kv in line.split() if "=" in kv)
            if parts.get("kernel") == kernel and "hbm_bytes" in parts:
                return float(parts["hbm_bytes"])
    return None


def measured_issue(workload: str, kernel: str):
    """VALU issue of `kernel` from the committed rocprofv3 SQ counter
    summary of this workload (SQ_ACTIVE_INST_VALU over the SIMDs' issue
    slots at the measured 0.474 wave64 instructions per SIMD per cycle), or
    None."""
    path = os.path.join(ROOT, "profiles", f"valu_issue_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = dict(kv.split("=", 1) for kv in line.split() if "=" in kv)
            if parts.get("kernel") == kernel and "valu_issue_frac" in parts:
                return {k: float(v) for k, v in parts.items() if k != "kernel"}
    return None


def cpu_baseline(sc, batch, threads: int, reps: int = 5, single_sample: int = 16):
    """The oracle chain over every event of the batch, one event per task on a
    pool of `threads` (ctypes releases the GIL): the median of `reps` runs in
    events/s; plus one thread over the first `single_sample` events.  Also
    returns the last run's per-event results (the checker of the GPU run)."""
    import concurrent.futures as cf
    import statistics

    from oracle import bridge
    from trax_amd import events as ev

    singles = [ev.concat_events([batch], subset=[e]) for e in range(batch.n_events)]
    runs, res = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            res = list(ex.map(lambda b: bridge.run_chain(sc, b), singles))
        runs.append(time.perf_counter() - t0)
    m = min(single_sample, len(singles))
    t0 = time.perf_counter()
    for b in singles[:m]:
        bridge.run_chain(sc, b)
    one = m / (time.perf_counter() - t0)
    dt = statistics.median(runs)
    return batch.n_events / dt, res, dt, one, m, runs


def measured_ta(workload: str, kernel: str):
    """Average-CU texture-address busy share of `kernel` from the committed
    TA counter summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", f"ta_busy_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 3 and cells[0] == kernel:
                try:
                    return float(cells[2]) / 100.0
                except ValueError:
                    return None
    return None


def limiter(kernel: str, issue, ta, traffic_frac):
    """What bounds the dominant kernel, in words, from its counters."""
    if kernel in ("predict_filter", "predict_walk", "filter_mask", "pairs"):
        txt = ("the mix of dependent gathers and VALU work, not HBM: per-lane window walks "
               "((z, phi) gathers after grid-cell loads, then compares) and the cuts' FP32 arithmetic")
        parts = []
        if issue:
            parts.append(f"VALU issue {100 * issue['valu_issue_frac']:.0f}% of the measured peak")
        if ta is not None:
            parts.append(f"texture-address units busy {100 * ta:.0f}%")
        if traffic_frac is not None:
            parts.append(f"HBM {100 * traffic_frac:.1f}% of 8 TB/s")
        return txt + ("; " + ", ".join(parts) if parts else "")
    return "HBM"


def check_against_oracle(got, batch, P: int, per_event) -> int:
    """Every event's triplets (as event-relative input-hit indices, per layer
    triplet) and segment sizes in the GPU batch result equal the oracle's on
    that event alone; the batch offsets are the running sums.  Raises on any
    difference; returns the number of triplets compared."""
    import numpy as np

    sizes = np.concatenate([np.diff(o["tracklet_offsets"].astype(np.int64)) for o in per_event])
    want_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    if not np.array_equal(got["tracklet_offsets"], want_off):
        raise SystemExit("triplet offsets differ from the oracle")

    def keys(rows, lt):
        r = rows.astype(np.uint64)
        return np.sort((lt.astype(np.uint64) << np.uint64(60)) | (r[:, 0] << np.uint64(40)) |
                       (r[:, 1] << np.uint64(20)) | r[:, 2])
    off, oi = got["tracklet_offsets"].astype(np.int64), got["orig_index"].astype(np.int64)
    for e, o in enumerate(per_event):
        base = int(batch.event_offset[e])
        seg = off[e * P:(e + 1) * P + 1]
        g_rows = oi[got["tracklets"][seg[0]:seg[-1]]] - base
        g_lt = np.repeat(np.arange(P), np.diff(seg))
        o_off = o["tracklet_offsets"].astype(np.int64)
        o_rows = o["orig_index"].astype(np.int64)[o["tracklets"]]
        o_lt = np.repeat(np.arange(P), np.diff(o_off))
        if not np.array_equal(keys(g_rows, g_lt), keys(o_rows, o_lt)):
            raise SystemExit(f"event {e}: triplet set differs from the oracle")
    return int(off[-1])


def real_events(args):
    """A batch of real events: the container's events in file order, cycled
    to --events (default: the file's count), selected and ordered as
    HitCollection::addEvent does, with the reference's ttbar configuration."""
    from trax_amd import cmsdata, pbio
    from trax_amd import events as ev
    if not args.geometry:
        raise SystemExit("--event-file needs --geometry")
    evs = pbio.load(args.event_file)
    n = args.events or len(evs)
    evs = [evs[i % len(evs)] for i in range(n)]
    g = cmsdata.Geometry.load(os.path.join(args.geometry, "detectorRadius.dat"),
                              os.path.join(args.geometry, "radiusDictionary.dat"))
    batch, _ = cmsdata.load_events(evs, g, cmsdata.reference_triplets(), 1.0)
    sc = cmsdata.ttbar_workload(g)
    w = ev.Workload(name=f"cms-{os.path.splitext(os.path.basename(args.event_file))[0]}-x{n}",
                    config_id=-1, seed_config=sc, n_events=n, radii=[])
    log(f"{n} real events from {args.event_file}")
    return w, batch


def connectivity_measure(ctx, runtime, bufs, ref, n_events: int, P: int, torch):
    """TripletConnectivity over the chain's triplets (downstream of the timed
    step): per layer triplet p the triplets of every event form one set, and
    tight connectivity runs p -> p+1 for p = 0..P-2, as the reference's
    multi-layer test chains run() calls.  Algorithmic bytes per call: hit2,
    hit3 and connectivity of each base triplet (12 B), hit1, hit2 and a
    connectivity read + write of each following triplet (16 B)."""
    import numpy as np
    if P < 2:
        return None
    off = ref["tracklet_offsets"].astype(np.int64)
    sets = []
    for p in range(P):
        idx = np.concatenate([np.arange(off[e * P + p], off[e * P + p + 1]) for e in range(n_events)])
        di = torch.from_numpy(idx).to("cuda")
        cols = [bufs.t[i].index_select(0, di) for i in range(3)]
        sets.append(runtime.Tracklets(cols[0], cols[1], cols[2], torch.zeros_like(cols[0])))
    ws = runtime.connectivity_workspace(ctx, max(t.n for t in sets[:-1]))

    def run_all():
        for p in range(P - 1):
            runtime.triplet_connectivity(ctx, sets[p], sets[p + 1], tight=True, workspace=ws)
    run_all()
    torch.cuda.synchronize()
    chains = [int((t.connectivity != 0).sum().item()) for t in sets[1:]]
    reps = 10
    ms, n = ctx.time_kernel("connect", run_all, reps)
    if ctx.flags():
        raise SystemExit("connectivity raised a device flag")
    nb = sum(sets[p].n for p in range(P - 1))
    nf = sum(sets[p + 1].n for p in range(P - 1))
    calls = n / reps
    b = (12 * nb + 16 * nf) / max(1.0, calls)
    return {"mode": "tight, layer triplet p -> p+1", "calls": int(calls), "base_triplets": nb,
            "following_triplets": nf, "following_with_chains": chains, "ms_per_call": ms,
            "bytes_per_call": b, "gbs": b / (ms * 1e-3) / 1e9}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described "
                                   "in DESIGN.md but not built", "n_gpus": args.gpus}))
        return 2
    import numpy as np
    import torch

    from trax_amd import events as ev
    from trax_amd import runtime

    torch.cuda.set_device(0)
    if args.profile:
        args.no_plain = args.no_cpu = True
    if args.event_file:
        w, batch = real_events(args)
        data = f"real ({os.path.basename(args.event_file)}, events cycled to the batch size)"
    else:
        w = ev.config_workload(args.config, args.events)
        log(f"generating {w.n_events} events ({w.name})")
        batch = ev.make_batch(w)
        data = "synthetic (seeded helix generator)"
    sc = w.seed_config
    sc.pair_generator = args.pair_generator
    if args.pair_generator != "beamspot":
        w.name += "-" + args.pair_generator
    log(f"{batch.n_hits} hits")
    ctx = runtime.Context(0)
    tune_path = args.tune
    if tune_path is None and not args.no_tune:
        p = os.path.join(ROOT, "trax_amd", "tuned", f"{w.name}.json")
        tune_path = p if os.path.exists(p) else None
    tuned = None
    if tune_path:
        from trax_amd import autotune
        with open(tune_path) as f:
            res = autotune.TuneResult(**json.load(f))
        autotune.apply(ctx, w, res)
        tuned = {"file": os.path.relpath(tune_path, ROOT), "launch": res.launch, "binning": f"{res.n_phi}phi x {res.n_z}z"}
        log(f"tuned: {res.n_phi} phi x {res.n_z} z, launch {res.launch}")
    bufs, db = runtime.run_batch(ctx, sc, batch)
    torch.cuda.synchronize()
    n_pairs, n_cands, n_trk = runtime.totals(bufs)
    log(f"pairs {n_pairs} candidates {n_cands} triplets {n_trk}")
    ref = runtime.fetch(bufs, batch.n_events, batch.n_hits)
    # the fused chain (no candidate records) on its own buffers; its results
    # must equal the staged chain's
    fbufs, fdb = runtime.run_batch(ctx, sc, batch, candidates=False)
    torch.cuda.synchronize()
    fref = runtime.fetch(fbufs, batch.n_events, batch.n_hits)
    keys = ("grid", "pairs", "pair_offsets", "cand_offsets", "tracklets", "tracklet_offsets")
    if runtime.totals(fbufs) != (n_pairs, n_cands, n_trk) or \
            not all(np.array_equal(fref[k], ref[k]) for k in keys):
        raise SystemExit("fused prediction + filter differs from the staged chain")
    if args.unfused:
        cbufs, cdb, cref, ckeys = bufs, db, ref, keys + ("cands",)
    else:
        cbufs, cdb, cref, ckeys = fbufs, fdb, fref, keys

    # ---- batched chain, captured once as a graph and replayed
    if args.no_graph:
        def step():
            runtime.seed_chain(ctx, cbufs, cdb)
    else:
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            runtime.seed_chain(ctx, cbufs, cdb)  # eager pass on the capture stream
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph):
            runtime.seed_chain(ctx, cbufs, cdb)
        torch.cuda.synchronize()

        def step():
            graph.replay()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(args.steps):
        step()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    flags = ctx.flags()
    got = runtime.fetch(cbufs, batch.n_events, batch.n_hits)
    same = all(np.array_equal(got[k], cref[k]) for k in ckeys)
    if flags or not same:
        raise SystemExit(f"replayed chain differs from the eager run (flags {flags}, same {same})")
    ev_s = batch.n_events / (ms / 1e3)
    log(f"batched: {ms:.3f} ms/step, {ev_s:.0f} events/s (wall {wall:.3f} s)")

    # ---- per-stage device time (eager, HIP events on the launch stream)
    st = {}
    E = db.n_events
    st["grid_build"] = time_events(lambda: runtime.grid_build(ctx, bufs, db), 5, torch)
    st["pairs"] = time_events(lambda: runtime.pair_stage(ctx, bufs, E), 5, torch)
    st["predict"] = time_events(lambda: runtime.triplet_predict(ctx, bufs, E), 5, torch)
    st["filter"] = time_events(lambda: runtime.triplet_filter(ctx, bufs, E), 5, torch)
    st["predict_filter"] = time_events(lambda: runtime.triplet_predict_filter(ctx, fbufs, E), 5, torch)
    log("stages ms: " + ", ".join(f"{k} {v:.3f}" for k, v in st.items()))

    # ---- per-kernel device time: the library brackets every launch of the
    # selected kernel with HIP events on its stream, over eager chain steps
    counts = {"hits": batch.n_hits, "pairs": n_pairs, "cands": n_cands, "triplets": n_trk,
              "cells": batch.n_events * sc.n_layers * sc.cells_per_layer}
    kern = {}
    reps = max(3, min(args.steps, 10))
    staged = ("predict_walk", "predict_place", "filter_mask", "filter_store")
    fused = ("predict_filter", "acc_cut", "triplet_place")
    for name in ("grid_build", "pairs") + staged + fused + ("scan",):
        b_, d_ = (fbufs, fdb) if name in fused else (bufs, db)
        ms_launch, n = ctx.time_kernel(name, lambda: runtime.seed_chain(ctx, b_, d_), reps)
        kern[name] = {"ms": ms_launch, "launches_per_step": n / reps}
        key = "pairs_sector" if name == "pairs" and sc.pair_generator == "sector" else name
        if key in ALGO_BYTES:
            b = ALGO_BYTES[key](counts) / max(1.0, n / reps)
            kern[name].update(bytes=b, gbs=b / (ms_launch * 1e-3) / 1e9)
    torch.cuda.synchronize()
    if ctx.flags():
        raise SystemExit("overflow during the per-kernel timing")
    log("kernels ms/launch: " + ", ".join(f"{k} {v['ms']:.3f}" for k, v in kern.items()))
    timed = [k for k in kern if k not in (fused if args.unfused else staged)]
    dominant = max(timed, key=lambda k: kern[k]["ms"] * kern[k]["launches_per_step"])

    conn = connectivity_measure(ctx, runtime, bufs, ref, batch.n_events, len(sc.triplets), torch)
    if conn:
        log(f"connectivity (tight, {conn['calls']} calls): {conn['ms_per_call']:.4f} ms/call, "
            f"{conn['gbs']:.0f} GB/s algorithmic")

    # scan roofline on a large array (4 B read + 4 B write per element), in
    # place, on seeded random values: an all-zero array scans about 12 %
    # faster on this part (zero lines are cheaper to move), which would
    # flatter the figure
    n_scan = 1 << 26
    g_scan = torch.Generator(device="cuda").manual_seed(26)
    big = torch.randint(-(1 << 31), (1 << 31) - 1, (n_scan,), dtype=torch.int32, device="cuda",
                        generator=g_scan)
    scan_ms, _ = ctx.time_kernel("scan", lambda: runtime.prefix_sum(ctx, big), 10)
    scan_gbs = 8 * n_scan / (scan_ms * 1e-3) / 1e9
    grid_gbs = kern["grid_build"]["gbs"]
    grid_stage_gbs = grid_stage_bytes(counts) / (kern["grid_build"]["ms"] * 1e-3) / 1e9
    log(f"grid build {grid_gbs:.0f} GB/s, scan(2^26) {scan_gbs:.0f} GB/s, dominant {dominant}")
    # elementwise kernel of known traffic (256 MiB read, 256 MiB written): the
    # measured copy ceiling of this device, and the calibration kernel of the
    # FETCH_SIZE / WRITE_SIZE counter runs
    src = torch.ones(64 << 20, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    copy_ms = time_events(lambda: torch.neg(src, out=dst), 10, torch)
    copy = {"kernel": "torch.neg on 2^26 int32", "bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4,
            "ms": copy_ms, "gbs": 2 * src.numel() * 4 / (copy_ms * 1e-3) / 1e9}
    copy["frac_of_peak"] = copy["gbs"] / HBM_PEAK_GBS
    log(f"copy 256 MiB: {copy_ms:.3f} ms, {copy['gbs']:.0f} GB/s")
    del src, dst

    # ---- PCIe: what a caller starting from host buffers adds (pinned copies of
    # the input columns and layout up, the triplet columns and offsets down)
    host_in = {k: t.cpu().pin_memory() for k, t in db.cols.items()}
    h2d_ms = time_events(lambda: [db.cols[k].copy_(host_in[k], non_blocking=True) for k in host_in],
                         20, torch)
    host_out = [torch.empty(n_trk, dtype=torch.int32, pin_memory=True) for _ in range(3)]
    host_off = torch.empty(bufs.tracklet_offsets.numel(), dtype=torch.int32, pin_memory=True)

    def d2h():
        for i in range(3):
            host_out[i].copy_(bufs.t[i][:n_trk], non_blocking=True)
        host_off.copy_(bufs.tracklet_offsets, non_blocking=True)
    d2h_ms = time_events(d2h, 20, torch)
    in_bytes = sum(t.numel() * t.element_size() for t in host_in.values())
    pcie = {"h2d_ms": h2d_ms, "d2h_ms": d2h_ms, "h2d_bytes": in_bytes,
            "d2h_bytes": 12 * n_trk + host_off.numel() * 4,
            "events_per_sec": batch.n_events / ((ms + h2d_ms + d2h_ms) * 1e-3)}
    log(f"pcie: up {h2d_ms:.3f} ms, down {d2h_ms:.3f} ms, inclusive {pcie['events_per_sec']:.0f} events/s")

    # ---- plain path: one event per launch sequence, host reads between stages
    plain = plain_c1 = None
    if not args.no_plain and not args.event_file:
        # the plain method on its own workload: pile-up-50 events, three
        # barrel layers, one event per launch sequence
        w1 = ev.config_workload(1, 64)
        b1 = ev.make_batch(w1)
        s1 = [runtime.DeviceBatch.upload(ev.concat_events([b1], subset=[e])) for e in range(b1.n_events)]
        p1, c1, t1 = runtime.default_capacities(max(d.n_hits for d in s1))
        pb1 = runtime.ChainBuffers(w1.seed_config, 1, max(d.n_hits for d in s1), p1, c1, t1)
        runtime.run_plain_event(ctx, pb1, s1[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n1 = sum(runtime.run_plain_event(ctx, pb1, d) for d in s1)
        torch.cuda.synchronize()
        dt1 = time.perf_counter() - t0
        plain_c1 = {"workload": w1.name, "events": len(s1), "hits": b1.n_hits, "triplets": n1,
                    "events_per_sec": len(s1) / dt1, "ms_per_event": dt1 * 1e3 / len(s1)}
        del pb1, s1
        log(f"plain ({w1.name}): {plain_c1['ms_per_event']:.3f} ms/event, {plain_c1['events_per_sec']:.0f} events/s")
    if not args.no_plain:
        singles = [runtime.DeviceBatch.upload(ev.concat_events([batch], subset=[e]))
                   for e in range(batch.n_events)]
        pb = runtime.ChainBuffers(sc, 1, max(d.n_hits for d in singles), bufs.pair_capacity,
                                  bufs.cand_capacity, bufs.tracklet_capacity)
        runtime.run_plain_event(ctx, pb, singles[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n_plain = 0
        for d in singles:
            n_plain += runtime.run_plain_event(ctx, pb, d)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        plain = {"events_per_sec": len(singles) / dt, "triplets": n_plain,
                 "ms_per_event": dt * 1e3 / len(singles)}
        if n_plain != n_trk:
            raise SystemExit(f"plain path found {n_plain} triplets, batched {n_trk}")
        log(f"plain: {plain['ms_per_event']:.3f} ms/event, {plain['events_per_sec']:.0f} events/s")

    # ---- CPU baseline (oracle, bounded sample = this batch)
    cpu = None
    if not args.no_cpu:
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = max(1, min(args.cpu_threads, avail))
        log(f"cpu baseline on {threads} threads ({avail} available)")
        cps, per_event, cdt, one, m, runs = cpu_baseline(sc, batch, threads)
        # the timed GPU step's triplets (the replayed graph's output) against
        # the oracle, event by event and segment by segment
        checked = check_against_oracle(got, batch, len(sc.triplets), per_event)
        log(f"oracle check: {checked} triplets of {batch.n_events} events equal")
        cpu = {"value": cps, "unit": "events/s", "cores": threads, "kind": "port",
               "sample": f"{batch.n_events} events of the same workload, C oracle chain, one event per task "
                         f"on {threads} threads, median of {len(runs)} runs ({cdt:.2f} s wall, runs "
                         f"{', '.join(f'{r:.2f}' for r in runs)} s)",
               "single_thread": {"value": one, "unit": "events/s", "cores": 1,
                                 "sample": f"the first {m} events on one thread"},
               "triplets_checked_vs_gpu": checked}

    dom = kern[dominant]
    traffic = measured_traffic(w.name, dominant)
    issue = measured_issue(w.name, dominant)
    ta_busy = measured_ta(w.name, dominant)
    out = {
        "metric": "events/sec (batched seeding chain, 1 GPU)",
        "value": ev_s,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u32",
        "data": data,
        "config": {"workload": w.name, "events": batch.n_events, "hits": batch.n_hits,
                   "layers": sc.n_layers, "layer_triplets": len(sc.triplets),
                   "grid": f"{sc.n_phi}phi x {sc.n_z}z", "graph": not args.no_graph,
                   "prediction_filter": "staged" if args.unfused else "fused", "tuned": tuned},
        "triplets_per_sec": n_trk / (ms / 1e3),
        "counts": {"pairs": n_pairs, "candidates": n_cands, "triplets": n_trk},
        "plain": plain,
        "batched_over_plain": (ev_s / plain["events_per_sec"]) if plain else None,
        "plain_config1": plain_c1,
        "pcie_inclusive": pcie,
        "stages_ms": st,
        "kernels": kern,
        "grid_build_roofline": {"achieved": grid_stage_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": grid_stage_gbs / HBM_PEAK_GBS,
                                "bytes": grid_stage_bytes(counts),
                                "achieved_min": grid_gbs, "frac_min": grid_gbs / HBM_PEAK_GBS,
                                "bytes_min": kern["grid_build"]["bytes"],
                                "traffic": measured_traffic(w.name, "grid_build"),
                                "ms": kern["grid_build"]["ms"]},
        "scan_roofline": {"n": n_scan, "achieved": scan_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": scan_gbs / HBM_PEAK_GBS, "ms": scan_ms},
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": f"profiles/hbm_traffic_{w.name}.txt (FETCH_SIZE + WRITE_SIZE runs of "
                                       f"bench.py --profile, copy-kernel calibrated)" if traffic else None,
                     "traffic_gbs": traffic / (dom["ms"] * 1e-3) / 1e9 if traffic else None,
                     "traffic_frac": traffic / (dom["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "ms": dom["ms"], "algorithmic_bytes": dom["bytes"],
                     "candidates_per_sec": (n_cands / (dom["ms"] * 1e-3)
                                            if dominant in ("predict_filter", "predict_walk") else None),
                     "limiter": limiter(dominant, issue, ta_busy, traffic / (dom["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
                                        if traffic else None),
                     "valu_issue": dict(issue, peak_wave_insts_per_simd_cycle=0.474,
                                        source=f"profiles/valu_issue_{w.name}.md") if issue else None,
                     "ta_busy": ta_busy},
        "cpu_baseline": cpu,
        "connectivity": conn,
        "copy_calibration": copy,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
