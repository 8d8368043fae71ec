"""ctypes bridge to the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and the CPU-baseline leg of
bench.py, where the oracle is the checker or the timed CPU baseline, never
the product path.  Builds nothing itself: `make -C oracle` (run by
__graft_entry__.build()) produces oracle/_build/liboracle.so.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TRAX_ORACLE_LIB selects another build of the same sources (the sanitized
# one of `make -C oracle sanitize`).
LIB_PATH = os.environ.get("TRAX_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")

MAX_LAYERS = 16
MAX_TRIPLETS = 8
UNWRITTEN = 0xFFFFFFFF

_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)


class OracleCfg(C.Structure):
    _fields_ = [
        ("n_events", C.c_uint32), ("n_layers", C.c_uint32),
        ("n_z", C.c_uint32), ("n_phi", C.c_uint32),
        ("min_z", C.c_float), ("sector_size_z", C.c_float),
        ("min_phi", C.c_float), ("sector_size_phi", C.c_float),
        ("min_radius", C.c_float * MAX_LAYERS), ("max_radius", C.c_float * MAX_LAYERS),
        ("n_triplets", C.c_uint32),
        ("layer1", C.c_uint32 * MAX_TRIPLETS), ("layer2", C.c_uint32 * MAX_TRIPLETS),
        ("layer3", C.c_uint32 * MAX_TRIPLETS),
        ("dtheta_cut", C.c_float * MAX_TRIPLETS), ("sigma_z", C.c_float * MAX_TRIPLETS),
        ("dphi_cut", C.c_float * MAX_TRIPLETS), ("sigma_phi", C.c_float * MAX_TRIPLETS),
        ("tip_cut", C.c_float * MAX_TRIPLETS), ("z0", C.c_float * MAX_TRIPLETS),
        ("pair_spread_z", C.c_uint32 * MAX_TRIPLETS), ("pair_spread_phi", C.c_uint32 * MAX_TRIPLETS),
        ("min_radius_curvature", C.c_float),
        ("threads", C.c_uint32),
        ("pair_generator", C.c_uint32),
    ]


class OracleInput(C.Structure):
    _fields_ = [
        ("n_hits", C.c_uint32),
        ("x", _f32p), ("y", _f32p), ("z", _f32p),
        ("layer", _u32p), ("det_id", _u32p), ("hit_id", _u32p), ("evt", _u32p),
        ("event_offset", _u32p), ("layer_nhits", _u32p), ("layer_offset", _u32p),
    ]


class OracleResult(C.Structure):
    _fields_ = [
        ("n_hits", C.c_uint32),
        ("x", _f32p), ("y", _f32p), ("z", _f32p),
        ("layer", _u32p), ("det_id", _u32p), ("hit_id", _u32p), ("evt", _u32p),
        ("orig_index", _u32p),
        ("grid_size", C.c_uint64), ("grid", _u32p),
        ("n_pairs", C.c_uint32), ("n_pairs_unwritten", C.c_uint32),
        ("n_pairs_raw_overcount", C.c_uint32),
        ("pairs", _u32p), ("pair_offsets", _u32p),
        ("n_cands", C.c_uint32), ("n_cands_unwritten", C.c_uint32),
        ("cands", _u32p), ("cand_offsets", _u32p),
        ("n_tracklets", C.c_uint32),
        ("t1", _u32p), ("t2", _u32p), ("t3", _u32p), ("tracklet_offsets", _u32p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.oracle_scan_u32.argtypes = [_u32p, C.c_uint64]
        L.oracle_run_chain.argtypes = [C.POINTER(OracleCfg), C.POINTER(OracleInput), C.c_int,
                                       C.POINTER(OracleResult)]
        L.oracle_run_chain.restype = C.c_int
        L.oracle_free_result.argtypes = [C.POINTER(OracleResult)]
        L.oracle_eval_f32.argtypes = [C.c_int, _f32p, _f32p, _f32p, C.c_uint64]
        L.oracle_cell_of.argtypes = [C.POINTER(OracleCfg), C.c_float, C.c_float, C.c_float, _f32p]
        L.oracle_cell_of.restype = C.c_uint32
        L.oracle_pair_window.argtypes = [C.POINTER(OracleCfg), C.c_uint32, C.c_float, C.c_float,
                                         C.c_float, _f32p]
        L.oracle_sector_window.argtypes = [C.POINTER(OracleCfg), C.c_uint32, C.c_float, C.c_float,
                                           C.c_float, _u32p]
        L.oracle_predict_window.argtypes = [C.POINTER(OracleCfg), C.c_uint32, C.c_int, _f32p, _f32p,
                                            C.c_float, C.c_float, _f32p]
        L.oracle_filter_eval.argtypes = [C.POINTER(OracleCfg), C.c_uint32, _f32p, _f32p, _f32p, _f32p]
        L.oracle_connectivity.argtypes = [_u32p, _u32p, _u32p, _u32p, C.c_uint64, _u32p, _u32p, _u32p, _u32p,
                                          C.c_uint64, C.c_int, C.c_int]
        L.oracle_connectivity.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


PAIR_GENERATORS = {"beamspot": 0, "sector": 1}  # ORACLE_PAIRS_BEAMSPOT / _SECTOR


def make_cfg(sc, n_events: int) -> OracleCfg:
    c = OracleCfg()
    c.n_events = n_events
    c.n_layers = sc.n_layers
    c.n_z, c.n_phi = sc.n_z, sc.n_phi
    c.min_z = float(np.float32(-300.0))
    c.sector_size_z = float(sc.sector_size_z)
    c.min_phi = float(np.float32(-np.pi))
    c.sector_size_phi = float(sc.sector_size_phi)
    for i in range(sc.n_layers):
        c.min_radius[i] = float(np.float32(sc.min_radius[i]))
        c.max_radius[i] = float(np.float32(sc.max_radius[i]))
    c.n_triplets = len(sc.triplets)
    for p, t in enumerate(sc.triplets):
        c.layer1[p], c.layer2[p], c.layer3[p] = t.layer1, t.layer2, t.layer3
        c.dtheta_cut[p] = float(np.float32(t.dtheta_cut))
        c.sigma_z[p] = float(np.float32(t.sigma_z))
        c.dphi_cut[p] = float(np.float32(t.dphi_cut))
        c.sigma_phi[p] = float(np.float32(t.sigma_phi))
        c.tip_cut[p] = float(np.float32(t.tip_cut))
        c.z0[p] = float(np.float32(int(t.z0)))
        c.pair_spread_z[p] = t.pair_spread_z
        c.pair_spread_phi[p] = t.pair_spread_phi
    c.min_radius_curvature = float(sc.min_radius_curvature)
    c.threads = sc.threads
    c.pair_generator = PAIR_GENERATORS[sc.pair_generator]
    return c


def _arr(ptr, n, dt):
    if n == 0 or not ptr:
        return np.zeros(n, dt)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)


def run_chain(sc, batch, stop_after: int = 3) -> Dict[str, np.ndarray]:
    """Run the oracle chain on a host EventBatch; returns numpy copies."""
    L = lib()
    cfg = make_cfg(sc, batch.n_events)
    keep = [np.ascontiguousarray(a) for a in batch.columns()]
    ev = np.ascontiguousarray(batch.event_offset, np.uint32)
    ln = np.ascontiguousarray(batch.layer_nhits, np.uint32)
    lo = np.ascontiguousarray(batch.layer_offset, np.uint32)
    inp = OracleInput(batch.n_hits, _p(keep[0], _f32p), _p(keep[1], _f32p), _p(keep[2], _f32p),
                      _p(keep[3], _u32p), _p(keep[4], _u32p), _p(keep[5], _u32p),
                      _p(keep[6], _u32p), _p(ev, _u32p), _p(ln, _u32p), _p(lo, _u32p))
    res = OracleResult()
    rc = L.oracle_run_chain(C.byref(cfg), C.byref(inp), stop_after, C.byref(res))
    if rc != 0:
        L.oracle_free_result(C.byref(res))
        raise RuntimeError(f"oracle_run_chain failed ({rc})")
    try:
        n = res.n_hits
        nk = batch.n_events * len(sc.triplets) + 1
        out = {
            "x": _arr(res.x, n, np.float32), "y": _arr(res.y, n, np.float32),
            "z": _arr(res.z, n, np.float32), "layer": _arr(res.layer, n, np.uint32),
            "det_id": _arr(res.det_id, n, np.uint32), "hit_id": _arr(res.hit_id, n, np.uint32),
            "evt": _arr(res.evt, n, np.uint32), "orig_index": _arr(res.orig_index, n, np.uint32),
            "grid": _arr(res.grid, res.grid_size, np.uint32),
        }
        if stop_after >= 1:
            out["pairs"] = _arr(res.pairs, 2 * res.n_pairs, np.uint32).reshape(-1, 2)
            out["pair_offsets"] = _arr(res.pair_offsets, nk, np.uint32)
            out["n_pairs_unwritten"] = int(res.n_pairs_unwritten)
            out["n_pairs_raw_overcount"] = int(res.n_pairs_raw_overcount)
        if stop_after >= 2:
            out["cands"] = _arr(res.cands, 2 * res.n_cands, np.uint32).reshape(-1, 2)
            out["cand_offsets"] = _arr(res.cand_offsets, nk, np.uint32)
            out["n_cands_unwritten"] = int(res.n_cands_unwritten)
        if stop_after >= 3:
            nt = res.n_tracklets
            out["tracklets"] = np.stack([_arr(res.t1, nt, np.uint32), _arr(res.t2, nt, np.uint32),
                                         _arr(res.t3, nt, np.uint32)], axis=1)
            out["tracklet_offsets"] = _arr(res.tracklet_offsets, nk, np.uint32)
        return out
    finally:
        L.oracle_free_result(C.byref(res))


def scan_u32(a: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(a, dtype=np.uint32).copy()
    lib().oracle_scan_u32(_p(b, _u32p), b.shape[0])
    return b


OPS = {"atan2": 0, "atan": 1, "acos": 2, "sign": 3, "div": 4, "sqrt": 5}


def eval_f32(op: str, a: np.ndarray, b: np.ndarray = None) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    b = a if b is None else np.ascontiguousarray(b, np.float32)
    out = np.empty_like(a)
    lib().oracle_eval_f32(OPS[op], _p(a, _f32p), _p(b, _f32p), _p(out, _f32p), a.shape[0])
    return out


WINDOW_FIELDS = ("zLow", "zHigh", "phiLow", "phiHigh", "wrap", "zLowSector", "zHighSector",
                 "phiLowSector", "phiHighSector")


def _window(vals):
    d = dict(zip(WINDOW_FIELDS, vals))
    for k in WINDOW_FIELDS[4:]:
        d[k] = int(d[k])
    return d


def pair_window(sc, p: int, x: float, y: float, z: float) -> dict:
    cfg = make_cfg(sc, 1)
    out = (C.c_float * 9)()
    lib().oracle_pair_window(C.byref(cfg), p, x, y, z, out)
    return _window(list(out))


def sector_window(sc, p: int, x: float, y: float, z: float) -> dict:
    """PairGeneratorSector window of one inner hit (rows and phi cells after
    the reference's unsigned / wrap-around corrections)."""
    cfg = make_cfg(sc, 1)
    out = (C.c_uint32 * 5)()
    lib().oracle_sector_window(C.byref(cfg), p, x, y, z, out)
    return {"zLowSector": out[0], "zHighSector": out[1], "phiLowSector": out[2],
            "phiHighSector": out[3], "wrap": bool(out[4])}


def predict_window(sc, lt: int, store: bool, p1, p2, min_r: float, max_r: float) -> dict:
    cfg = make_cfg(sc, 1)
    a = (C.c_float * 3)(*[float(v) for v in p1])
    b = (C.c_float * 3)(*[float(v) for v in p2])
    out = (C.c_float * 9)()
    lib().oracle_predict_window(C.byref(cfg), lt, int(store), a, b, min_r, max_r, out)
    return _window(list(out))


def filter_eval(sc, lt: int, h1, h2, h3) -> dict:
    cfg = make_cfg(sc, 1)
    arrs = [(C.c_float * 3)(*[float(v) for v in h]) for h in (h1, h2, h3)]
    out = (C.c_float * 5)()
    lib().oracle_filter_eval(C.byref(cfg), lt, arrs[0], arrs[1], arrs[2], out)
    return {"theta_ratio": out[0], "dphi": out[1], "tip2": out[2], "radius": out[3],
            "pass": bool(out[4])}


def cell_of(sc, x: float, y: float, z: float):
    cfg = make_cfg(sc, 1)
    phi = C.c_float()
    cell = lib().oracle_cell_of(C.byref(cfg), x, y, z, C.byref(phi))
    return int(cell), phi.value


def connectivity(base, following, backward: bool = False, tight: bool = False) -> np.ndarray:
    """TripletConnectivity::run on host arrays.  base / following: (hit1, hit2,
    hit3, connectivity) uint32 arrays; returns the following set's updated
    connectivity (the inputs are not modified)."""
    L = lib()
    b = [np.ascontiguousarray(a, np.uint32) for a in base]
    f = [np.ascontiguousarray(a, np.uint32) for a in following]
    out = f[3].copy()
    rc = L.oracle_connectivity(_p(b[0], _u32p), _p(b[1], _u32p), _p(b[2], _u32p), _p(b[3], _u32p), len(b[3]),
                               _p(f[0], _u32p), _p(f[1], _u32p), _p(f[2], _u32p), _p(out, _u32p), len(out),
                               int(backward), int(tight))
    if rc != 0:
        raise RuntimeError(f"oracle_connectivity failed ({rc})")
    return out
