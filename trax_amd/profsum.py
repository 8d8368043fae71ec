"""Summaries of rocprofv3 output for the seeding kernels.

`stats`  : per-kernel table (calls, average / total device time, share, and
           the launch shape and register / LDS use from the kernel trace) of
           a `rocprofv3 --kernel-trace --stats --output-format csv` run.
`sq`     : instruction-issue view of each kernel from one counter run with
           SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VMEM_RD SQ_WAVE_CYCLES SQ_BUSY_CYCLES
           SQ_WAIT_INST_ANY SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU
           GRBM_GUI_ACTIVE: VALU instructions per wave, the share of the
           SIMDs' measured peak VALU issue rate used (VALU_PEAK_PER_SIMD_CYCLE),
           the share of lanes active per VALU instruction, and the share of
           wave time spent waiting on dependencies.
`ta`     : texture-address unit busy share and SALU / LDS instructions per
           wave from one counter run with TA_BUSY_avr TA_BUSY_max
           GRBM_GUI_ACTIVE SQ_INSTS_SALU SQ_INSTS_LDS SQ_WAVES.
`counters`: per-kernel averages of every counter of one counter run, e.g.
           `--pmc OccupancyPercent LdsBankConflict LdsUtil SQ_WAVES` (wave
           occupancy, LDS bank-conflict cycles per LDS access, LDS busy %).
`traffic`: HBM bytes per launch of each chain kernel from two counter runs
           (`--pmc FETCH_SIZE`, then `--pmc WRITE_SIZE`), each counter scaled
           by what it reads for a copy kernel of known size (torch's int32
           negation of a 256 MiB array: 256 MiB read, 256 MiB written).

Both print text meant to be kept with the command that produced the data:

    python -m trax_amd.profsum stats DIR --command "..."
    python -m trax_amd.profsum traffic FETCH_DIR WRITE_DIR --workload NAME --command "..."
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re
from collections import defaultdict
from typing import Dict, List

# Chain kernels by the short names the library's kernel timer uses.
SHORT = [
    ("k_grid_cells", "grid_cells"),  # earlier builds
    ("k_grid_sort", "grid_sort"),
    ("k_pairs_count", "pairs_count"),
    ("k_pairs_store", "pairs_store"),
    ("k_sector_pairs<false>", "sector_count"),
    ("k_sector_pairs<true>", "sector_store"),
    ("k_predict_walk", "predict_walk"),
    ("k_predict_filter_walk", "predict_filter"),
    ("k_acc_redo", "acc_redo"),
    ("k_acc_cut", "acc_cut"),
    ("k_acc_place", "triplet_place"),
    ("k_acc_offsets", "acc_offsets"),  # earlier builds
    ("k_fused_offsets", "fused_offsets"),
    ("k_predict_place", "predict_place"),
    ("k_predict<", "predict_two_phase"),
    ("k_filter_mask", "filter_mask"),
    ("k_filter_store", "filter_store"),
    ("k_scan_u32", "scan"),
    ("k_scan_reduce", "scan_reduce"),
    ("k_scan_down", "scan_down"),
    ("k_gather_offsets", "gather_offsets"),
    ("k_pred_offsets", "pred_offsets"),
    ("k_pair_plan", "pair_plan"),
    ("k_pair_counts", "pair_counts"),
    ("k_conn_clear", "connect_clear"),
    ("k_conn_insert", "connect_insert"),
    ("k_conn_lookup", "connect_lookup"),
    ("k_read_last", "read_last"),
]
COPY_BYTES = 256 << 20  # calibration kernel: bytes read and bytes written


def short_name(name: str) -> str:
    for key, short in SHORT:
        if key in name:
            return short
    if "neg_kernel" in name:
        return "copy_calibration"
    return re.sub(r"\(.*", "", name)[:60]


def _one(pattern: str) -> str:
    files = sorted(glob.glob(pattern, recursive=True))
    if not files:
        raise SystemExit(f"no file matches {pattern}")
    return files[0]


def _rows(path: str) -> List[dict]:
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def stats(d: str) -> str:
    st = _rows(_one(os.path.join(d, "**", "*kernel_stats.csv")))
    shape: Dict[str, dict] = {}
    tr = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)
    if tr:
        for r in _rows(tr[0]):
            n = short_name(r["Kernel_Name"])
            shape.setdefault(n, {"vgpr": r.get("VGPR_Count"), "agpr": r.get("Accum_VGPR_Count"),
                                 "sgpr": r.get("SGPR_Count"), "lds": r.get("LDS_Block_Size"),
                                 "scratch": r.get("Scratch_Size"), "wg": r.get("Workgroup_Size_X"),
                                 "grid": r.get("Grid_Size_X")})
    out = ["| kernel | calls | avg us | total ms | % | VGPR | AGPR | SGPR | LDS B | scratch | WG | grid |",
           "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for r in st:
        n = short_name(r["Name"])
        s = shape.get(n, {})
        out.append(f"| {n} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} | "
                   f"{s.get('vgpr', '')} | {s.get('agpr', '')} | {s.get('sgpr', '')} | {s.get('lds', '')} | "
                   f"{s.get('scratch', '')} | {s.get('wg', '')} | {s.get('grid', '')} |")
    return "\n".join(out)


def _counter_means(d: str, counter: str) -> Dict[str, float]:
    rows = _rows(_one(os.path.join(d, "**", "*counter_collection.csv")))
    per: Dict[str, List[float]] = defaultdict(list)
    for r in rows:
        if r.get("Counter_Name", counter) != counter:
            continue
        per[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in per.items()}


# Wave64 FP32 VALU instructions one SIMD issues per GRBM_GUI_ACTIVE cycle at
# most, measured on MI355X with a kernel of independent multiplies (32 per
# loop iteration, four independent chains, 8192 blocks of 256 threads):
# 2.1e9 wave instructions in 2.03 ms over 1024 SIMDs and 4.32 M GUI cycles
# per XCD, i.e. one instruction per SIMD about every 2.1 cycles.
VALU_PEAK_PER_SIMD_CYCLE = 0.474


def sq(d: str, xcds: int = 8, simds: int = 1024, valu_rate: float = VALU_PEAK_PER_SIMD_CYCLE) -> str:
    rows = _rows(_one(os.path.join(d, "**", "*counter_collection.csv")))
    per: Dict[str, Dict[str, List[float]]] = defaultdict(lambda: defaultdict(list))
    for r in rows:
        per[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = ["| kernel | waves | VALU insts / wave | VMEM reads / wave | VALU issue % of peak | active lanes % | "
           "wait % |", "|---|---|---|---|---|---|---|"]
    machine = []
    for k, cs in per.items():
        c = {n: sum(v) / len(v) for n, v in cs.items()}
        if not c.get("SQ_WAVES") or not c.get("GRBM_GUI_ACTIVE") or not c.get("SQ_WAVE_CYCLES"):
            continue
        cycles = c["GRBM_GUI_ACTIVE"] / xcds  # GRBM_GUI_ACTIVE sums the XCDs
        # SQ_ACTIVE_INST_VALU counts issued VALU instructions (it equals
        # SQ_INSTS_VALU on the probe kernel), so the share of the peak issue
        # rate is instructions / (SIMDs x cycles x peak rate)
        issue = c["SQ_ACTIVE_INST_VALU"] / (simds * cycles * valu_rate)
        lanes = c["SQ_THREAD_CYCLES_VALU"] / (64.0 * c["SQ_ACTIVE_INST_VALU"]) if c["SQ_ACTIVE_INST_VALU"] else 0
        out.append(f"| {k} | {c['SQ_WAVES']:.0f} | {c['SQ_INSTS_VALU'] / c['SQ_WAVES']:.0f} | "
                   f"{c['SQ_INSTS_VMEM_RD'] / c['SQ_WAVES']:.0f} | {100 * issue:.0f} | {100 * lanes:.0f} | "
                   f"{100 * c['SQ_WAIT_INST_ANY'] / c['SQ_WAVE_CYCLES']:.0f} |")
        machine.append(f"kernel={k} waves={c['SQ_WAVES']:.0f} valu_insts={c['SQ_INSTS_VALU']:.0f} "
                       f"active_inst_valu={c['SQ_ACTIVE_INST_VALU']:.0f} gui_cycles={cycles:.0f} "
                       f"valu_issue_frac={issue:.4f} active_lanes_frac={lanes:.4f}")
    return "\n".join(out + [""] + machine)


def ta(d: str, xcds: int = 8, tas: int = 256) -> str:
    """Texture-address unit busy share and SALU / LDS instructions per wave
    from one counter run with TA_BUSY_avr TA_BUSY_max GRBM_GUI_ACTIVE
    SQ_INSTS_SALU SQ_INSTS_LDS SQ_WAVES (TA_BUSY_avr: busy cycles averaged
    over the TA instances, one per CU)."""
    rows = _rows(_one(os.path.join(d, "**", "*counter_collection.csv")))
    per: Dict[str, Dict[str, List[float]]] = defaultdict(lambda: defaultdict(list))
    for r in rows:
        per[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = ["| kernel | waves | TA busy % (avg CU) | TA busy % (busiest CU) | SALU insts / wave | LDS insts / wave |",
           "|---|---|---|---|---|---|"]
    for k, cs in per.items():
        c = {n: sum(v) / len(v) for n, v in cs.items()}
        if not c.get("SQ_WAVES") or not c.get("GRBM_GUI_ACTIVE"):
            continue
        cycles = c["GRBM_GUI_ACTIVE"] / xcds
        out.append(f"| {k} | {c['SQ_WAVES']:.0f} | {100 * c.get('TA_BUSY_avr', 0) / cycles:.0f} | "
                   f"{100 * c.get('TA_BUSY_max', 0) / cycles:.0f} | {c.get('SQ_INSTS_SALU', 0) / c['SQ_WAVES']:.0f} | "
                   f"{c.get('SQ_INSTS_LDS', 0) / c['SQ_WAVES']:.0f} |")
    return "\n".join(out)


def counters(d: str, kernels=None) -> str:
    """Per-kernel average of every counter in one counter-collection run
    (raw counters and rocprofv3's derived metrics alike), one column each."""
    rows = _rows(_one(os.path.join(d, "**", "*counter_collection.csv")))
    per: Dict[str, Dict[str, List[float]]] = defaultdict(lambda: defaultdict(list))
    names: List[str] = []
    for r in rows:
        k = short_name(r["Kernel_Name"])
        if kernels and k not in kernels:
            continue
        n = r["Counter_Name"]
        if n not in names:
            names.append(n)
        per[k][n].append(float(r["Counter_Value"]))
    out = ["| kernel | " + " | ".join(names) + " |", "|---|" + "---|" * len(names)]
    for k, cs in per.items():
        vals = []
        for n in names:
            v = cs.get(n)
            vals.append(f"{sum(v) / len(v):.3g}" if v else "")
        out.append(f"| {k} | " + " | ".join(vals) + " |")
    return "\n".join(out)


def traffic(fetch_dir: str, write_dir: str, workload: str) -> str:
    fetch = _counter_means(fetch_dir, "FETCH_SIZE")
    write = _counter_means(write_dir, "WRITE_SIZE")
    if "copy_calibration" not in fetch or "copy_calibration" not in write:
        raise SystemExit("the calibration copy kernel is missing from the counter runs")
    kib = COPY_BYTES / 1024.0
    f_scale = kib / fetch["copy_calibration"]
    w_scale = kib / write["copy_calibration"]
    lines = [f"# workload={workload} fetch_scale={f_scale:.4f} write_scale={w_scale:.4f} "
             f"(copy kernel: FETCH_SIZE {fetch['copy_calibration']:.0f} KiB, WRITE_SIZE "
             f"{write['copy_calibration']:.0f} KiB for {kib:.0f} KiB each way)"]
    total = {}
    for k in sorted(set(fetch) & set(write)):
        fb = fetch[k] * f_scale * 1024.0
        wb = write[k] * w_scale * 1024.0
        total[k] = fb + wb
        lines.append(f"kernel={k} fetch_kib={fetch[k]:.1f} write_kib={write[k]:.1f} "
                     f"hbm_bytes={fb + wb:.0f}")
    # the grid stage is the per-segment sort (earlier builds ran a separate
    # cell kernel before it)
    if "grid_sort" in total:
        parts = [k for k in ("grid_cells", "grid_sort") if k in total]
        lines.append(f"kernel=grid_build hbm_bytes={sum(total[k] for k in parts):.0f} "
                     f"({' + '.join(parts)})")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("--command", default="")
    c = sub.add_parser("sq")
    c.add_argument("dir")
    c.add_argument("--command", default="")
    t = sub.add_parser("ta")
    t.add_argument("dir")
    t.add_argument("--command", default="")
    d = sub.add_parser("counters")
    d.add_argument("dir")
    d.add_argument("--kernels", default="", help="comma-separated short names (default: all)")
    d.add_argument("--command", default="")
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--workload", required=True)
    b.add_argument("--command", default="")
    args = ap.parse_args()
    if args.command:
        print(f"# command: {args.command}")
    if args.what == "stats":
        print(stats(args.dir))
    elif args.what == "sq":
        print(sq(args.dir))
    elif args.what == "ta":
        print(ta(args.dir))
    elif args.what == "counters":
        print(counters(args.dir, [k for k in args.kernels.split(",") if k] or None))
    else:
        print(traffic(args.fetch_dir, args.write_dir, args.workload))


if __name__ == "__main__":
    main()
