"""The device's FP32 built-in definitions against the oracle's, on the CPU.

trax_amd/csrc/fp32_ocl.h (what the kernels call) writes the range reductions
of atan2, atan and acos with selects, so that a divergent wave evaluates one
division, square root and polynomial; oracle/ofp32.h keeps the plain branch
form.  Both use only correctly rounded IEEE binary32 operations, so compiled
for the host (SSE, no contraction) they must agree bit for bit; the GPU test
test_fp32_definitions_bitwise checks the device build against the oracle."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_device_fp32_definitions_equal_oracle_on_host(tmp_path):
    exe = str(tmp_path / "fp32_defs")
    src = os.path.join(ROOT, "tests", "fp32_defs_vs_oracle.c")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-msse2", "-mfpmath=sse",
                    "-I" + os.path.join(ROOT, "trax_amd", "csrc"), "-I" + os.path.join(ROOT, "oracle"),
                    src, "-o", exe, "-lm"], check=True)
    r = subprocess.run([exe, "5000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("0 mismatches")
