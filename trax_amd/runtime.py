"""Python host side over the C ABI of libtraxhip.so (include/trax_hip.h).

Device buffers are torch tensors on the current HIP device (PyTorch is used
only for device memory and streams); every computation runs in the gfx950
kernels behind the C ABI.  If the library is missing the import of this
module still works, but creating a Context raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from . import events as ev

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libtraxhip.so")
# Kernel ids of the context's kernel timer (TRAX_KERNEL_* in include/trax_hip.h).
KERNELS = {"grid_build": 1, "pairs": 2, "predict_walk": 3, "predict_place": 4, "filter_mask": 5,
           "filter_store": 6, "scan": 7, "connect": 8, "predict_filter": 9, "triplet_place": 10,
           "acc_cut": 11}
# Store-window hits per pair the single-call predictor buffers before a pair walks again.
PREDICT_BUFFER = 16
PREDICT_TILE, PREDICT_TILE_SLOTS = 64, 512  # TRAX_PREDICT_TILE, TRAX_PREDICT_TILE_SLOTS

MAX_LAYERS = 16
MAX_TRIPLETS = 8
UNWRITTEN = 0xFFFFFFFF
FLAG_PAIR_OVERFLOW = 0x1
FLAG_CAND_OVERFLOW = 0x2
FLAG_TRACKLET_OVERFLOW = 0x4
FLAG_LAYOUT = 0x8

_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)
_vp = C.c_void_p


class LayerTripletC(C.Structure):
    _fields_ = [("layer1", C.c_uint32), ("layer2", C.c_uint32), ("layer3", C.c_uint32),
                ("dtheta_cut", C.c_float), ("sigma_z", C.c_float), ("dphi_cut", C.c_float),
                ("sigma_phi", C.c_float), ("tip_cut", C.c_float), ("pair_spread_z", C.c_uint32),
                ("z0", C.c_float), ("pair_spread_phi", C.c_uint32), ("n_candidates", C.c_uint32)]


class SeedConfigC(C.Structure):
    _fields_ = [("n_z", C.c_uint32), ("n_phi", C.c_uint32), ("n_layers", C.c_uint32),
                ("min_radius", C.c_float * MAX_LAYERS), ("max_radius", C.c_float * MAX_LAYERS),
                ("n_triplets", C.c_uint32), ("triplets", LayerTripletC * MAX_TRIPLETS),
                ("min_radius_curvature", C.c_float), ("pair_generator", C.c_uint32)]


PAIR_GENERATORS = {"beamspot": 0, "sector": 1}  # TRAX_PAIRS_BEAMSPOT / TRAX_PAIRS_SECTOR


class HitsC(C.Structure):
    _fields_ = [("x", _vp), ("y", _vp), ("z", _vp), ("layer", _vp), ("det_id", _vp),
                ("hit_id", _vp), ("evt", _vp), ("orig_index", _vp), ("phi", _vp), ("zphi", _vp),
                ("xyzphi", _vp)]


class LayoutC(C.Structure):
    _fields_ = [("n_events", C.c_uint32), ("n_hits", C.c_uint32), ("event_offset", _vp),
                ("layer_nhits", _vp), ("layer_offset", _vp)]


class WorkspaceC(C.Structure):
    _fields_ = [("work_offsets", _vp), ("pair_counts", _vp), ("cand_counts", _vp),
                ("max_events", C.c_uint32), ("max_hits", C.c_uint32),
                ("pair_capacity", C.c_uint64), ("cand_capacity", C.c_uint64),
                ("cand_stash", _vp), ("filter_mask", _vp), ("filter_counts", _vp),
                ("cand_tile_counts", _vp), ("trk_tile_counts", _vp),
                ("n_triplets", C.c_uint32), ("pair_items", C.c_uint32)]


class TrackletsC(C.Structure):
    _fields_ = [("hit1", C.c_void_p), ("hit2", C.c_void_p), ("hit3", C.c_void_p),
                ("connectivity", C.c_void_p), ("n", C.c_uint64)]


class LaunchC(C.Structure):
    _fields_ = [("grid_waves", C.c_uint32), ("pairs_block", C.c_uint32),
                ("predict_block", C.c_uint32), ("scan_tile", C.c_uint32),
                ("fused_buffer", C.c_uint32), ("fused_block_waves", C.c_uint32),
                ("fused_waves_per_eu", C.c_uint32)]


# Compiled (lane buffer, waves per block, waves per SIMD) shapes of the
# fused prediction + filter walk (trax_launch_config); the first is the default.
FUSED_SHAPES = ((8, 4, 7), (8, 4, 6), (8, 4, 8), (4, 4, 8), (4, 4, 7), (8, 2, 7), (12, 4, 6))


class OutputsC(C.Structure):
    _fields_ = [("pairs", _vp), ("pair_capacity", C.c_uint64), ("pair_offsets", _vp),
                ("cands", _vp), ("cand_capacity", C.c_uint64), ("cand_offsets", _vp),
                ("t1", _vp), ("t2", _vp), ("t3", _vp), ("tracklet_capacity", C.c_uint64),
                ("tracklet_offsets", _vp), ("totals", _vp)]


# Every symbol include/trax_hip.h declares, with its ctypes signature.
_P = C.POINTER
SIGNATURES = {
    "trax_abi_version": (C.c_int, []),
    "trax_ctx_create": (C.c_int, [C.c_int, _P(_vp)]),
    "trax_ctx_destroy": (C.c_int, [_vp]),
    "trax_ctx_last_error": (C.c_char_p, [_vp]),
    "trax_ctx_flags": (C.c_int, [_vp, _vp, _u32p, C.c_int]),
    "trax_ctx_set_launch": (C.c_int, [_vp, _P(LaunchC)]),
    "trax_ctx_get_launch": (C.c_int, [_vp, _P(LaunchC)]),
    "trax_ctx_kernel_timer": (C.c_int, [_vp, C.c_int]),
    "trax_ctx_kernel_time": (C.c_int, [_vp, _P(C.c_double), _u32p]),
    "trax_min_radius_curvature": (C.c_float, [C.c_float]),
    "trax_workspace_layout": (C.c_uint64, [_P(SeedConfigC), C.c_uint32, C.c_uint32, C.c_uint64,
                                           C.c_uint64, _vp, _P(WorkspaceC)]),
    "trax_prefix_sum": (C.c_int, [_vp, _vp, C.c_uint64, _vp]),
    "trax_connectivity_workspace_bytes": (C.c_uint64, [C.c_uint64]),
    "trax_triplet_connectivity": (C.c_int, [_vp, _P(TrackletsC), _P(TrackletsC), C.c_uint32, _vp, C.c_uint64,
                                            _vp]),
    "trax_prefix_sum_dev": (C.c_int, [_vp, _vp, _vp, C.c_uint64, _vp, C.c_uint32, _vp]),
    "trax_grid_build": (C.c_int, [_vp, _P(SeedConfigC), _P(HitsC), _P(LayoutC), _P(HitsC), _vp,
                                  _vp]),
    "trax_pairs_count": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp, _P(WorkspaceC),
                                   _P(OutputsC), _vp]),
    "trax_pairs_store": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp, _P(WorkspaceC),
                                   _P(OutputsC), _vp]),
    "trax_pairs_beamspot": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp,
                                      _P(WorkspaceC), _P(OutputsC), _vp]),
    "trax_pairs_sector_count": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp,
                                          _P(WorkspaceC), _P(OutputsC), _vp]),
    "trax_pairs_sector_store": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp,
                                          _P(WorkspaceC), _P(OutputsC), _vp]),
    "trax_pairs_sector": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp,
                                    _P(WorkspaceC), _P(OutputsC), _vp]),
    "trax_triplet_predict": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp,
                                       _P(WorkspaceC), _P(OutputsC), _vp]),
    "trax_predict_count": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp,
                                     _P(WorkspaceC), _P(OutputsC), _vp]),
    "trax_predict_store": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp,
                                     _P(WorkspaceC), _P(OutputsC), _vp]),
    "trax_filter_count": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _P(WorkspaceC),
                                    _P(OutputsC), _vp]),
    "trax_filter_store": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(WorkspaceC), _P(OutputsC),
                                    _vp]),
    "trax_triplet_filter": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _P(WorkspaceC),
                                      _P(OutputsC), _vp]),
    "trax_triplet_predict_filter": (C.c_int, [_vp, _P(SeedConfigC), C.c_uint32, _P(HitsC), _vp,
                                              _P(WorkspaceC), _P(OutputsC), _vp]),
    "trax_seed_chain": (C.c_int, [_vp, _P(SeedConfigC), _P(HitsC), _P(LayoutC), _P(HitsC), _vp,
                                  _P(WorkspaceC), _P(OutputsC), _vp]),
    "trax_prefix_sum_segmented": (C.c_int, [_vp, _vp, C.c_uint64, _vp, C.c_uint32, _vp]),
    "trax_prefix_sum_noalloc": (C.c_int, [_vp, C.c_uint64, _vp]),
    "trax_chain_graph_capture": (C.c_int, [_vp, _P(SeedConfigC), _P(HitsC), _P(LayoutC), _P(HitsC), _vp,
                                           _P(WorkspaceC), _P(OutputsC), _P(_vp)]),
    "trax_chain_graph_launch": (C.c_int, [_vp, _vp]),
    "trax_chain_graph_destroy": (C.c_int, [_vp]),
    "trax_fp32_eval": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_uint64, _vp]),
}

_lib = None


def lib():
    """Load libtraxhip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m trax_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def seed_config_c(sc: ev.SeedConfig) -> SeedConfigC:
    c = SeedConfigC()
    c.n_z, c.n_phi, c.n_layers = sc.n_z, sc.n_phi, sc.n_layers
    for i in range(sc.n_layers):
        c.min_radius[i] = float(np.float32(sc.min_radius[i]))
        c.max_radius[i] = float(np.float32(sc.max_radius[i]))
    c.n_triplets = len(sc.triplets)
    for p, t in enumerate(sc.triplets):
        r = c.triplets[p]
        r.layer1, r.layer2, r.layer3 = t.layer1, t.layer2, t.layer3
        r.dtheta_cut = float(np.float32(t.dtheta_cut))
        r.sigma_z = float(np.float32(t.sigma_z))
        r.dphi_cut = float(np.float32(t.dphi_cut))
        r.sigma_phi = float(np.float32(t.sigma_phi))
        r.tip_cut = float(np.float32(t.tip_cut))
        r.pair_spread_z = t.pair_spread_z
        r.z0 = float(np.float32(int(t.z0)))
        r.pair_spread_phi = t.pair_spread_phi
        r.n_candidates = t.n_candidates
    c.min_radius_curvature = float(sc.min_radius_curvature)
    c.pair_generator = PAIR_GENERATORS[sc.pair_generator]
    return c


def _torch():
    import torch
    return torch


class TraxError(RuntimeError):
    pass


class Context:
    """One trax context (scan control block, status flags) on one device."""

    def __init__(self, device: int = 0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise TraxError("no HIP device: the seeding path has no CPU fallback")
        self.lib = lib()
        self.device = device
        torch.cuda.set_device(device)
        h = _vp()
        rc = self.lib.trax_ctx_create(device, C.byref(h))
        if rc != 0:
            raise TraxError(f"trax_ctx_create failed ({rc})")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.trax_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.trax_ctx_last_error(self.h)
            raise TraxError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_launch(self, grid_waves=0, pairs_block=256, predict_block=256, scan_tile=4096,
                   fused_buffer=0, fused_block_waves=0, fused_waves_per_eu=0):
        cfg = LaunchC(grid_waves, pairs_block, predict_block, scan_tile, fused_buffer, fused_block_waves,
                      fused_waves_per_eu)
        self.check(self.lib.trax_ctx_set_launch(self.h, C.byref(cfg)), "trax_ctx_set_launch")

    def get_launch(self) -> dict:
        cfg = LaunchC()
        self.check(self.lib.trax_ctx_get_launch(self.h, C.byref(cfg)), "trax_ctx_get_launch")
        return {k: getattr(cfg, k) for k, _ in LaunchC._fields_}

    def time_kernel(self, kernel: str, fn, reps: int):
        """Average device time (ms) of one launch of `kernel` over `reps`
        eager calls of fn(), from HIP events the library records around each
        launch on its stream; returns (ms_per_launch, launches)."""
        kid = KERNELS[kernel]
        self.check(self.lib.trax_ctx_kernel_timer(self.h, kid), "trax_ctx_kernel_timer")
        try:
            for _ in range(reps):
                fn()
            ms, n = C.c_double(0.0), C.c_uint32(0)
            self.check(self.lib.trax_ctx_kernel_time(self.h, C.byref(ms), C.byref(n)),
                       "trax_ctx_kernel_time")
        finally:
            self.lib.trax_ctx_kernel_timer(self.h, 0)
        return (ms.value / n.value if n.value else float("nan")), int(n.value)

    def flags(self, stream=None, clear: bool = True) -> int:
        v = C.c_uint32(0)
        self.check(self.lib.trax_ctx_flags(self.h, stream_ptr(stream), C.byref(v), int(clear)),
                   "trax_ctx_flags")
        return int(v.value)


def stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return _vp(s.cuda_stream)


def _ptr(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _dev_u32(a: np.ndarray, device):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a).astype(np.uint32).view(np.int32)).to(device)


def _dev_f32(a: np.ndarray, device):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device)


def _empty(n: int, kind: str, device):
    torch = _torch()
    dt = torch.float32 if kind == "f" else torch.int32
    return torch.empty(max(int(n), 1), dtype=dt, device=device)


@dataclass
class DeviceBatch:
    """An EventBatch resident on the device."""

    n_events: int
    n_hits: int
    cols: Dict[str, object]

    @staticmethod
    def upload(batch: ev.EventBatch, device="cuda") -> "DeviceBatch":
        cols = {
            "x": _dev_f32(batch.x, device), "y": _dev_f32(batch.y, device),
            "z": _dev_f32(batch.z, device), "layer": _dev_u32(batch.layer, device),
            "det_id": _dev_u32(batch.det_id, device), "hit_id": _dev_u32(batch.hit_id, device),
            "evt": _dev_u32(batch.evt, device),
            "event_offset": _dev_u32(batch.event_offset, device),
            "layer_nhits": _dev_u32(batch.layer_nhits, device),
            "layer_offset": _dev_u32(batch.layer_offset, device),
        }
        return DeviceBatch(batch.n_events, batch.n_hits, cols)

    def hits_c(self) -> HitsC:
        c = self.cols
        return HitsC(_ptr(c["x"]), _ptr(c["y"]), _ptr(c["z"]), _ptr(c["layer"]), _ptr(c["det_id"]),
                     _ptr(c["hit_id"]), _ptr(c["evt"]), None, None)

    def layout_c(self) -> LayoutC:
        c = self.cols
        return LayoutC(self.n_events, self.n_hits, _ptr(c["event_offset"]), _ptr(c["layer_nhits"]),
                       _ptr(c["layer_offset"]))


class ChainBuffers:
    """Resident device buffers of one chain: sorted hits, grid, workspace, outputs.
    With candidates=False no candidate records are kept: the chain then runs
    the fused prediction + filter pass (the same triplets and offsets)."""

    def __init__(self, sc: ev.SeedConfig, max_events: int, max_hits: int, pair_capacity: int,
                 cand_capacity: int, tracklet_capacity: int, device="cuda", candidates: bool = True):
        torch = _torch()
        self.sc = sc
        self.cfg = seed_config_c(sc)
        self.max_events, self.max_hits = int(max_events), int(max_hits)
        self.pair_capacity = int(pair_capacity)
        self.candidates = bool(candidates)
        self.cand_capacity = int(cand_capacity) if candidates else 0
        self.tracklet_capacity = int(tracklet_capacity)
        n = self.max_hits
        self.sorted = {k: _empty(n, "f" if k in ("x", "y", "z", "phi") else "u", device)
                       for k in ("x", "y", "z", "layer", "det_id", "hit_id", "evt", "orig_index", "phi")}
        self.sorted["zphi"] = _empty(2 * n, "f", device)
        self.sorted["xyzphi"] = _empty(4 * n, "f", device)
        K1 = self.max_events * len(sc.triplets) + 1
        self.grid = _empty(self.max_events * sc.n_layers * sc.cells_per_layer, "u", device)
        L = lib()
        ws = WorkspaceC()
        nbytes = L.trax_workspace_layout(C.byref(self.cfg), self.max_events, self.max_hits,
                                         self.pair_capacity, self.cand_capacity, None, C.byref(ws))
        self.ws_mem = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
        self.ws = WorkspaceC()
        L.trax_workspace_layout(C.byref(self.cfg), self.max_events, self.max_hits, self.pair_capacity,
                                self.cand_capacity, _vp(self.ws_mem.data_ptr()), C.byref(self.ws))
        self.pairs = _empty(2 * self.pair_capacity, "u", device)
        self.pair_offsets = _empty(K1, "u", device)
        self.cands = _empty(2 * self.cand_capacity, "u", device) if candidates else None
        self.cand_offsets = _empty(K1, "u", device)
        self.t = [_empty(self.tracklet_capacity, "u", device) for _ in range(3)]
        self.tracklet_offsets = _empty(K1, "u", device)
        self.totals = torch.zeros(4, dtype=torch.int32, device=device)

    def sorted_c(self) -> HitsC:
        s = self.sorted
        return HitsC(*[_ptr(s[k]) for k in ("x", "y", "z", "layer", "det_id", "hit_id", "evt",
                                            "orig_index", "phi", "zphi", "xyzphi")])

    def outputs_c(self) -> OutputsC:
        return OutputsC(_ptr(self.pairs), self.pair_capacity, _ptr(self.pair_offsets),
                        _ptr(self.cands) if self.candidates else None, self.cand_capacity,
                        _ptr(self.cand_offsets),
                        _ptr(self.t[0]), _ptr(self.t[1]), _ptr(self.t[2]), self.tracklet_capacity,
                        _ptr(self.tracklet_offsets), _ptr(self.totals))

    def fits(self, n_events: int, n_hits: int) -> bool:
        return n_events <= self.max_events and n_hits <= self.max_hits


def default_capacities(n_hits: int):
    """Generous first guesses (pairs ~ 40 per hit at pile-up 200); grown on overflow."""
    pairs = max(1024, 64 * n_hits)
    return pairs, 2 * pairs, max(1024, pairs // 4)


def seed_chain(ctx: Context, bufs: ChainBuffers, db: DeviceBatch, stream=None) -> None:
    """Enqueue the whole chain (no host synchronisation)."""
    cfg = bufs.cfg
    hin, lay, hs, out = db.hits_c(), db.layout_c(), bufs.sorted_c(), bufs.outputs_c()
    bufs._keep = (hin, lay, hs, out)
    ctx.check(ctx.lib.trax_seed_chain(ctx.h, C.byref(cfg), C.byref(hin), C.byref(lay), C.byref(hs),
                                      _vp(_ptr(bufs.grid)), C.byref(bufs.ws), C.byref(out),
                                      stream_ptr(stream)), "trax_seed_chain")


def totals(bufs: ChainBuffers):
    t = bufs.totals.cpu().numpy().view(np.uint32)
    return int(t[0]), int(t[1]), int(t[2])


def run_batch(ctx: Context, sc: ev.SeedConfig, batch: ev.EventBatch, device="cuda",
              bufs: Optional[ChainBuffers] = None, max_retries: int = 4, candidates: bool = True):
    """Run the chain on a batch, growing the buffers until nothing overflows.
    candidates=False runs the fused prediction + filter (no candidate
    records).  Returns (bufs, db)."""
    db = DeviceBatch.upload(batch, device)
    if bufs is None or not bufs.fits(db.n_events, db.n_hits):
        p, c, t = default_capacities(db.n_hits)
        bufs = ChainBuffers(sc, db.n_events, max(db.n_hits, 1), p, c, t, device, candidates)
    for _ in range(max_retries):
        ctx.flags(clear=True)
        seed_chain(ctx, bufs, db)
        f = ctx.flags(clear=True)
        if f & FLAG_LAYOUT:
            raise TraxError("inconsistent event layout")
        if not (f & (FLAG_PAIR_OVERFLOW | FLAG_CAND_OVERFLOW | FLAG_TRACKLET_OVERFLOW)):
            return bufs, db
        n_p, n_c, n_t = totals(bufs)
        bufs = ChainBuffers(sc, bufs.max_events, bufs.max_hits, max(n_p, bufs.pair_capacity),
                            max(n_c, bufs.cand_capacity), max(n_t, bufs.tracklet_capacity), device,
                            bufs.candidates)
    raise TraxError("buffers still overflow after regrowing")


def fetch(bufs: ChainBuffers, n_events: int, n_hits: int) -> Dict[str, np.ndarray]:
    """Copy the chain's results to the host (numpy)."""
    n_p, n_c, n_t = totals(bufs)
    K1 = n_events * len(bufs.sc.triplets) + 1
    G = n_events * bufs.sc.n_layers * bufs.sc.cells_per_layer

    def u(t, n):
        return t[:n].cpu().numpy().view(np.uint32).copy()

    def f(t, n):
        return t[:n].cpu().numpy().copy()

    s = bufs.sorted
    return {
        "x": f(s["x"], n_hits), "y": f(s["y"], n_hits), "z": f(s["z"], n_hits),
        "phi": f(s["phi"], n_hits), "layer": u(s["layer"], n_hits), "det_id": u(s["det_id"], n_hits),
        "hit_id": u(s["hit_id"], n_hits), "evt": u(s["evt"], n_hits),
        "orig_index": u(s["orig_index"], n_hits), "grid": u(bufs.grid, G),
        "pairs": u(bufs.pairs, 2 * n_p).reshape(-1, 2), "pair_offsets": u(bufs.pair_offsets, K1),
        "cands": u(bufs.cands, 2 * n_c).reshape(-1, 2) if bufs.candidates else None,
        "cand_offsets": u(bufs.cand_offsets, K1),
        "tracklets": np.stack([u(bufs.t[i], n_t) for i in range(3)], axis=1),
        "tracklet_offsets": u(bufs.tracklet_offsets, K1),
    }


def fetch_triplets(bufs: ChainBuffers, n_events: int, n_hits: int) -> Dict[str, np.ndarray]:
    """Copy only the chain's result to the host: the triplets (as sorted-hit
    indexes), their per-(event, layer triplet) offsets and the sorted-to-input
    index map."""
    n_t = totals(bufs)[2]
    K1 = n_events * len(bufs.sc.triplets) + 1

    def u(t, n):
        return t[:n].cpu().numpy().view(np.uint32).copy()

    return {"orig_index": u(bufs.sorted["orig_index"], n_hits),
            "tracklets": np.stack([u(bufs.t[i], n_t) for i in range(3)], axis=1),
            "tracklet_offsets": u(bufs.tracklet_offsets, K1)}


CONNECT_BACKWARD, CONNECT_TIGHT = 0x1, 0x2  # TRAX_CONNECT_*


@dataclass
class Tracklets:
    """A device triplet set: hit columns and the connectivity column
    (int32 tensors holding uint32 values); columns a mode does not read may
    be None."""

    hit1: Optional[object]
    hit2: Optional[object]
    hit3: Optional[object]
    connectivity: object

    @property
    def n(self) -> int:
        return int(self.connectivity.numel())

    def c(self) -> TrackletsC:
        return TrackletsC(_ptr(self.hit1), _ptr(self.hit2), _ptr(self.hit3), _ptr(self.connectivity), self.n)


def connectivity_workspace(ctx: Context, n_base: int, device="cuda"):
    """A workspace tensor for trax_triplet_connectivity over n_base base triplets."""
    torch = _torch()
    nbytes = int(ctx.lib.trax_connectivity_workspace_bytes(n_base))
    return torch.empty(max(nbytes // 4, 1), dtype=torch.int32, device=device)


def triplet_connectivity(ctx: Context, base: Tracklets, following: Tracklets, backward: bool = False,
                         tight: bool = False, workspace=None, stream=None) -> None:
    """TripletConnectivity::run(base, following, iterateBackwards, tightPacking):
    adds to following.connectivity in place."""
    ws = workspace if workspace is not None else connectivity_workspace(ctx, base.n, following.connectivity.device)
    mode = (CONNECT_BACKWARD if backward else 0) | (CONNECT_TIGHT if tight else 0)
    b, f = base.c(), following.c()
    ctx.check(ctx.lib.trax_triplet_connectivity(ctx.h, C.byref(b), C.byref(f), mode, _vp(ws.data_ptr()),
                                                ws.numel() * 4, stream_ptr(stream)),
              "trax_triplet_connectivity")


def prefix_sum(ctx: Context, t, stream=None) -> None:
    """In-place exclusive scan of a 16-byte aligned int32/uint32 device tensor."""
    ctx.check(ctx.lib.trax_prefix_sum(ctx.h, _vp(t.data_ptr()), t.numel(), stream_ptr(stream)),
              "trax_prefix_sum")


def prefix_sum_segmented(ctx: Context, t, seg_offsets, stream=None) -> None:
    """In-place exclusive scan of a device tensor that restarts at every
    segment start; seg_offsets: device int32 tensor of n_segments + 1 offsets
    (0 first, t.numel() last, non-decreasing)."""
    n_seg = int(seg_offsets.numel()) - 1
    ctx.check(ctx.lib.trax_prefix_sum_segmented(ctx.h, _vp(t.data_ptr()), t.numel(),
                                                _vp(seg_offsets.data_ptr()), max(n_seg, 0),
                                                stream_ptr(stream)), "trax_prefix_sum_segmented")


class ChainGraph:
    """The whole chain over one resident batch recorded once as a HIP graph by
    the library (trax_chain_graph_capture) and replayed with launch()."""

    def __init__(self, ctx: Context, bufs: ChainBuffers, db: DeviceBatch):
        self.ctx = ctx
        cfg = bufs.cfg
        self._keep = (cfg, db.hits_c(), db.layout_c(), bufs.sorted_c(), bufs.outputs_c(), bufs, db)
        _, hin, lay, hs, out, _, _ = self._keep
        h = _vp()
        ctx.check(ctx.lib.trax_chain_graph_capture(ctx.h, C.byref(cfg), C.byref(hin), C.byref(lay), C.byref(hs),
                                                   _vp(_ptr(bufs.grid)), C.byref(bufs.ws), C.byref(out),
                                                   C.byref(h)), "trax_chain_graph_capture")
        self.h = h

    def launch(self, stream=None) -> None:
        ctx = self.ctx
        ctx.check(ctx.lib.trax_chain_graph_launch(self.h, stream_ptr(stream)), "trax_chain_graph_launch")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.ctx.lib.trax_chain_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


OPS = {"atan2": 0, "atan": 1, "acos": 2, "sign": 3, "div": 4, "sqrt": 5}


def fp32_eval(ctx: Context, op: str, a, b=None, stream=None):
    torch = _torch()
    b = a if b is None else b
    out = torch.empty_like(a)
    ctx.check(ctx.lib.trax_fp32_eval(ctx.h, OPS[op], _vp(a.data_ptr()), _vp(b.data_ptr()),
                                     _vp(out.data_ptr()), a.numel(), stream_ptr(stream)),
              "trax_fp32_eval")
    return out


# ------------------------------------------------------------ stage-by-stage
def _stage_args(bufs: ChainBuffers):
    return bufs.cfg, bufs.sorted_c(), bufs.outputs_c()


def grid_build(ctx: Context, bufs: ChainBuffers, db: DeviceBatch, stream=None) -> None:
    """GridBuilder::run (GridBuilder.h:42) on the device batch."""
    hin, lay, hs = db.hits_c(), db.layout_c(), bufs.sorted_c()
    ctx.check(ctx.lib.trax_grid_build(ctx.h, C.byref(bufs.cfg), C.byref(hin), C.byref(lay),
                                      C.byref(hs), _vp(_ptr(bufs.grid)), stream_ptr(stream)),
              "trax_grid_build")


def _stage(ctx, fn, name, bufs, n_events, with_grid=True, stream=None):
    cfg, hs, out = _stage_args(bufs)
    bufs._keep = (hs, out)
    if with_grid:
        rc = fn(ctx.h, C.byref(cfg), n_events, C.byref(hs), _vp(_ptr(bufs.grid)), C.byref(bufs.ws),
                C.byref(out), stream_ptr(stream))
    else:
        rc = fn(ctx.h, C.byref(cfg), n_events, C.byref(hs), C.byref(bufs.ws), C.byref(out),
                stream_ptr(stream))
    ctx.check(rc, name)


def pairs_count(ctx, bufs, n_events, stream=None):
    """Count phase of the configured pair generator (beam-spot or sector)."""
    if bufs.sc.pair_generator == "sector":
        _stage(ctx, ctx.lib.trax_pairs_sector_count, "trax_pairs_sector_count", bufs, n_events,
               stream=stream)
    else:
        _stage(ctx, ctx.lib.trax_pairs_count, "trax_pairs_count", bufs, n_events, stream=stream)


def pairs_store(ctx, bufs, n_events, stream=None):
    """Store phase of the configured pair generator."""
    if bufs.sc.pair_generator == "sector":
        _stage(ctx, ctx.lib.trax_pairs_sector_store, "trax_pairs_sector_store", bufs, n_events,
               stream=stream)
    else:
        _stage(ctx, ctx.lib.trax_pairs_store, "trax_pairs_store", bufs, n_events, stream=stream)


def predict_count(ctx, bufs, n_events, stream=None):
    _stage(ctx, ctx.lib.trax_predict_count, "trax_predict_count", bufs, n_events, stream=stream)


def predict_store(ctx, bufs, n_events, stream=None):
    _stage(ctx, ctx.lib.trax_predict_store, "trax_predict_store", bufs, n_events, stream=stream)


def pairs_beamspot(ctx, bufs, n_events, stream=None):
    """Single-pass pair stage (count, positions and store in one kernel)."""
    _stage(ctx, ctx.lib.trax_pairs_beamspot, "trax_pairs_beamspot", bufs, n_events, stream=stream)


def pairs_sector(ctx, bufs, n_events, stream=None):
    """PairGeneratorSector (PairGeneratorSector.h:38-39): count, scan and store
    with no host read."""
    _stage(ctx, ctx.lib.trax_pairs_sector, "trax_pairs_sector", bufs, n_events, stream=stream)


def pair_stage(ctx, bufs, n_events, stream=None):
    """Single-call pair stage of the configured generator."""
    if bufs.sc.pair_generator == "sector":
        pairs_sector(ctx, bufs, n_events, stream)
    else:
        pairs_beamspot(ctx, bufs, n_events, stream)


def triplet_predict(ctx, bufs, n_events, stream=None):
    """Single-call prediction stage (walk + stash, scan, compaction; no host read)."""
    _stage(ctx, ctx.lib.trax_triplet_predict, "trax_triplet_predict", bufs, n_events, stream=stream)


def triplet_predict_filter(ctx, bufs, n_events, stream=None):
    """Fused prediction + filter (no candidate records; no host read)."""
    _stage(ctx, ctx.lib.trax_triplet_predict_filter, "trax_triplet_predict_filter", bufs, n_events,
           stream=stream)


def filter_count(ctx, bufs, n_events, stream=None):
    _stage(ctx, ctx.lib.trax_filter_count, "trax_filter_count", bufs, n_events, with_grid=False,
           stream=stream)


def filter_store(ctx, bufs, n_events, stream=None):
    cfg, _, out = _stage_args(bufs)
    bufs._keep = (out,)
    ctx.check(ctx.lib.trax_filter_store(ctx.h, C.byref(cfg), n_events, C.byref(bufs.ws), C.byref(out),
                                        stream_ptr(stream)), "trax_filter_store")


def triplet_filter(ctx, bufs, n_events, stream=None):
    _stage(ctx, ctx.lib.trax_triplet_filter, "trax_triplet_filter", bufs, n_events,
           with_grid=False, stream=stream)


def run_plain_event(ctx: Context, bufs: ChainBuffers, db: DeviceBatch) -> int:
    """The plain method for one resident event: every stage is its own set of
    launches and the host reads each stage's total before the next store, as
    the reference does (PairGeneratorBeamspot.cpp:120, TripletThetaPhi-
    Predictor.cpp:59, TripletThetaPhiFilter.cpp:82).  Returns the triplet count."""
    torch = _torch()
    host = torch.empty(4, dtype=torch.int32, pin_memory=True)
    grid_build(ctx, bufs, db)
    pairs_count(ctx, bufs, db.n_events)
    host.copy_(bufs.totals)  # blocking read of the pair total
    if int(host[0]) > bufs.pair_capacity:
        raise TraxError("pair capacity exceeded in the plain path")
    pairs_store(ctx, bufs, db.n_events)
    predict_count(ctx, bufs, db.n_events)
    host.copy_(bufs.totals)  # blocking read of the candidate total
    if int(host[1]) > bufs.cand_capacity:
        raise TraxError("candidate capacity exceeded in the plain path")
    predict_store(ctx, bufs, db.n_events)
    filter_count(ctx, bufs, db.n_events)
    host.copy_(bufs.totals)  # blocking read of the triplet total
    if int(host[2]) > bufs.tracklet_capacity:
        raise TraxError("triplet capacity exceeded in the plain path")
    filter_store(ctx, bufs, db.n_events)
    return int(host[2])
