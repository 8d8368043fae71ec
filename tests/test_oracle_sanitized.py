"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: its C
restatement, built with -fsanitize=address,undefined and
-fno-sanitize-recover=all (`make -C oracle sanitize`), runs every entry
point the other oracle tests use (chains of all five workloads including
multi-event batches, the sector generator, connectivity, the scan, the FP32
built-ins and the window/filter probes) in a child process with the
sanitizer runtimes preloaded.  Any out-of-bounds access, use after free,
signed overflow, invalid shift or misaligned load aborts the child."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import numpy as np
from oracle import bridge
from trax_amd import events as ev
for cid, n in ((0, 1), (1, 2), (2, 4), (3, 2), (4, 3)):
    w = ev.config_workload(cid, n)
    b = ev.make_batch(w)
    o = bridge.run_chain(w.seed_config, b)
    assert len(o["tracklets"]) > 0, cid
    for T in (1, 7):
        w.seed_config.threads = T
        bridge.run_chain(w.seed_config, b, stop_after=1)
    w.seed_config.threads = 256
w = ev.config_workload(1, 2)
w.seed_config.pair_generator = "sector"
bridge.run_chain(w.seed_config, ev.make_batch(w))
w = ev.config_workload(1, 1)
for t in w.seed_config.triplets:
    t.sigma_phi, t.sigma_z = 3.0, 1.0  # wide windows: long candidate lists
bridge.run_chain(w.seed_config, ev.make_batch(w))
rng = np.random.default_rng(1)
a = rng.integers(0, 1 << 32, 100003, dtype=np.uint64).astype(np.uint32)
bridge.scan_u32(a)
x = rng.standard_normal(10000).astype(np.float32)
for op in ("atan2", "atan", "acos", "sign", "div", "sqrt"):
    bridge.eval_f32(op, np.clip(x, -1, 1) if op == "acos" else x, x[::-1].copy())
sc = ev.config_workload(3, 1).seed_config
bridge.pair_window(sc, 0, 10.0, -3.0, 5.0)
bridge.sector_window(sc, 0, 4.0, 1.0, 2.0)
bridge.predict_window(sc, 0, True, (4.0, 0.5, 1.0), (7.0, 1.0, 2.0), 10.0, 10.4)
bridge.filter_eval(sc, 0, (4.0, 0.1, 1.0), (7.0, 0.3, 1.8), (10.0, 0.6, 2.6))
cols = [rng.integers(0, 50, 5000).astype(np.uint32) for _ in range(4)]
for back, tight in ((False, False), (True, False), (False, True)):
    bridge.connectivity(cols, cols, backward=back, tight=tight)
maps = open("/proc/self/maps").read()
assert "liboracle_san.so" in maps and "libasan" in maps, "the sanitized build was not loaded"
print("sanitized oracle ok")
'''


def test_oracle_under_address_and_ub_sanitizers():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitized oracle does not build here: " + r.stderr[-400:])
    libs = []
    for name in ("libasan.so", "libubsan.so"):
        p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
        if not os.path.isabs(p) or not os.path.exists(p):
            pytest.skip(f"{name} not found")
        libs.append(p)
    env = dict(os.environ)
    env.update(LD_PRELOAD=":".join(libs), ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TRAX_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "liboracle_san.so"),
               PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", CHILD], capture_output=True, text=True, env=env, timeout=600)
    report = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0 and "sanitized oracle ok" in r.stdout, report
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, report
