"""Build libtraxhip.so (the gfx950 kernels + C ABI) in-tree with hipcc.

Every HIP source is compiled for gfx950 with IEEE FP32 semantics: no FMA
contraction (-ffp-contract=off), no fast-math, correctly rounded division
and square root, denormals kept, and packed-FP32 instructions disabled.
The output lands in trax_amd/_build/, next to the sources it comes from.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libtraxhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TRAX_OFFLOAD_ARCH", "gfx950")

DRIVERS = ["trax_seed", "trax_connect"]
SOURCES = ["scan.hip", "grid.hip", "pairs.hip", "sector.hip", "predict.hip", "filter.hip", "connect.hip", "capi.hip",
           "noalloc.hip"]

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if err.strip():
        sys.stderr.write(err + "\n")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    # C++ drivers over the algorithm classes (host code only, link the library)
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    for name in DRIVERS:
        drv = os.path.join(BUILD, name)
        cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
               "-I" + os.path.join(rocm, "include"), "-I" + os.path.join(ROOT, "include"),
               os.path.join(CSRC, "host", name + ".cpp"), "-L" + BUILD, "-ltraxhip",
               "-L" + os.path.join(rocm, "lib"), "-lamdhip64",
               "-Wl,-rpath,$ORIGIN:" + os.path.join(rocm, "lib"), "-o", drv + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name} build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(drv + ".tmp", drv)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
