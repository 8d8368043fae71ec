"""Binary event-group files read by the C++ driver trax_seed (format in
trax_amd/csrc/host/trax_seed.cpp), and its results files."""
from __future__ import annotations

import struct

import numpy as np

from . import events as ev


def write_events(path: str, sc: ev.SeedConfig, batch: ev.EventBatch) -> None:
    L, P = sc.n_layers, len(sc.triplets)
    with open(path, "wb") as f:
        f.write(struct.pack("<4sIIIIIIIf", b"TRXE", 1, batch.n_events, L, batch.n_hits, sc.n_z,
                            sc.n_phi, P, sc.min_pt))
        f.write(np.asarray(sc.min_radius[:L], np.float32).tobytes())
        f.write(np.asarray(sc.max_radius[:L], np.float32).tobytes())
        for t in sc.triplets:
            f.write(struct.pack("<IIIfffffIfII", t.layer1, t.layer2, t.layer3, t.dtheta_cut,
                                t.sigma_z, t.dphi_cut, t.sigma_phi, t.tip_cut, t.pair_spread_z,
                                float(int(t.z0)), t.pair_spread_phi, t.n_candidates))
        for a in (batch.event_offset, batch.layer_nhits, batch.layer_offset):
            f.write(np.asarray(a, np.uint32).tobytes())
        for a in batch.columns():
            f.write(np.ascontiguousarray(a).tobytes())


def read_results(path: str, n_events: int, n_triplets: int) -> dict:
    K1 = n_events * n_triplets + 1
    buf = open(path, "rb").read()
    pos = 0

    def u32(n):
        nonlocal pos
        a = np.frombuffer(buf, np.uint32, n, pos)
        pos += 4 * n
        return a.copy()

    assert buf[:4] == b"TRXR"
    pos = 4
    n = int(u32(1)[0])
    out = {"orig_index": u32(n)}
    g = int(u32(1)[0])
    out["grid"] = u32(g)
    n_p = int(u32(1)[0])
    out["pairs"] = u32(2 * n_p).reshape(-1, 2)
    out["pair_offsets"] = u32(K1)
    n_c = int(u32(1)[0])
    out["cands"] = u32(2 * n_c).reshape(-1, 2)
    out["cand_offsets"] = u32(K1)
    n_t = int(u32(1)[0])
    t1, t2, t3 = u32(n_t), u32(n_t), u32(n_t)
    out["tracklets"] = np.stack([t1, t2, t3], axis=1)
    out["tracklet_offsets"] = u32(K1)
    return out
