"""The C ABI library loads and exports every symbol include/trax_hip.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from trax_amd import events as ev
from trax_amd import runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "trax_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(trax_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_abi():
    names = _declared()
    assert "trax_seed_chain" in names and "trax_prefix_sum" in names
    assert set(names) == set(runtime.SIGNATURES), set(names) ^ set(runtime.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = runtime.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_host_only_entry_points():
    lib = runtime.lib()
    assert lib.trax_abi_version() == 8
    # connectivity table: a power of two >= 2 * n_base slots (at least 64), plus one, 16 B each
    for n, slots in ((0, 64), (1, 64), (32, 64), (33, 128), (1000, 2048), (1 << 20, 1 << 21)):
        assert lib.trax_connectivity_workspace_bytes(n) == (slots + 1) * 16
    for pt in (0.3, 0.5, 1.0):
        assert np.float32(lib.trax_min_radius_curvature(pt)) == ev.min_radius_curvature(pt)
    # 1 GeV at Bz = 3.8112 T: 87.52 cm (TripletConfiguration.cpp:145-152)
    assert abs(lib.trax_min_radius_curvature(1.0) - 87.522) < 1e-3


def test_workspace_layout_sizes():
    lib = runtime.lib()
    sc = ev.config_workload(2, 4).seed_config
    cfg = runtime.seed_config_c(sc)
    ws = runtime.WorkspaceC()
    n = lib.trax_workspace_layout(ctypes.byref(cfg), 4, 10000, 50000, 100000, None, ctypes.byref(ws))
    words = (100000 + 63) // 64 + 8
    need = ((4 * 2 + 2) * 4 + 20001 * 4 + 50001 * 4   # work offsets, pair (2 per hit) and candidate counts
            + (50000 + 63) // 64 * 512 * 4            # predictor stash slabs
            + words * 8 + (words + 1) * 4             # filter mask words and their counts
            + 2 * ((50000 + 63) // 64 + 2) * 4)       # candidate slots and triplets per tile of pairs
    assert need <= n <= need + 8 * 256
    assert ws.max_events == 4 and ws.max_hits == 10000 and ws.pair_capacity == 50000
    assert ws.cand_capacity == 100000
    # the layout records what it was sized for: the stage calls check against it
    assert ws.n_triplets == 2 and ws.pair_items == 2 * 10000
    base = 1 << 20
    lib.trax_workspace_layout(ctypes.byref(cfg), 4, 10000, 50000, 100000, ctypes.c_void_p(base),
                              ctypes.byref(ws))
    ptrs = [ws.work_offsets, ws.pair_counts, ws.cand_counts, ws.cand_stash, ws.filter_mask,
            ws.filter_counts, ws.cand_tile_counts, ws.trk_tile_counts]
    for p in ptrs:
        assert p % 256 == 0 and base <= p < base + n
    # regions in order and disjoint
    sizes = [(4 * 2 + 2) * 4, 20001 * 4, 50001 * 4, (50000 + 63) // 64 * 512 * 4, words * 8,
             (words + 1) * 4, ((50000 + 63) // 64 + 2) * 4, ((50000 + 63) // 64 + 2) * 4]
    for (p, sz), q in zip(zip(ptrs, sizes), ptrs[1:] + [base + n]):
        assert p + sz <= q
    # the grid build needs no workspace, so the binning does not change it
    big = ev.config_workload(2, 4).seed_config
    big.n_z, big.n_phi = 32, 64
    cfgb = runtime.seed_config_c(big)
    nb = lib.trax_workspace_layout(ctypes.byref(cfgb), 4, 10000, 50000, 100000, ctypes.c_void_p(base),
                                   ctypes.byref(ws))
    assert nb == n


def test_context_requires_a_device():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(runtime.TraxError):
        runtime.Context(0)


def test_workspace_pair_items_cover_shared_middle_layers():
    """Layer triplets that share a middle layer each walk its hits once in
    the beam-spot pair passes: the pair counts are sized for the largest
    such group (triplets (1,3,4), (2,3,5), (3,4,5): middle layer 3 twice)."""
    lib = runtime.lib()
    sc = ev.config_workload(3, 1).seed_config
    sc.triplets = [ev.LayerTriplet(1, 3, 4), ev.LayerTriplet(2, 3, 5), ev.LayerTriplet(3, 4, 5)]
    cfg = runtime.seed_config_c(sc)
    ws = runtime.WorkspaceC()
    n = lib.trax_workspace_layout(ctypes.byref(cfg), 2, 5000, 1000, 1000, None, ctypes.byref(ws))
    assert n > 0 and ws.n_triplets == 3 and ws.pair_items == 2 * 2 * 5000
    # distinct middle layers: one walk per hit
    sc.triplets = [ev.LayerTriplet(1, 2, 3), ev.LayerTriplet(2, 3, 4), ev.LayerTriplet(3, 4, 5)]
    cfg = runtime.seed_config_c(sc)
    lib.trax_workspace_layout(ctypes.byref(cfg), 2, 5000, 1000, 1000, None, ctypes.byref(ws))
    assert ws.pair_items == 2 * 5000
