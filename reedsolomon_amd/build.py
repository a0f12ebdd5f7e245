"""Builds reedsolomon_amd/librsx.so (HIP for gfx950) in-tree with hipcc.

    python reedsolomon_amd/build.py

Every .hip / .cpp file under csrc/ is compiled to an object for gfx950 and
linked into one shared library exporting the C ABI of include/rsx.h.  The
HBM streaming microbenchmark that bench.py uses for the device's read/write
ceilings (membench/) is built beside it into membench/libmembench.so.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librsx.so")
MEMBENCH_SRC = [os.path.join(HERE, "membench", f) for f in ("membench.hip", "butterfly_rates.hip", "hostloops.hip")]
MEMBENCH_LIB = os.path.join(HERE, "membench", "libmembench.so")
ARCH = os.environ.get("RSX_OFFLOAD_ARCH", "gfx950")
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Wall", "-Wno-unused-result",
    # no packed-FP32 instruction selection anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           f"-L{ROCM_LIB}", "-lrocprofiler-sdk-roctx"]  # trace ranges (trace.hpp)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    tmp = MEMBENCH_LIB + ".tmp"
    cmd = [HIPCC, *COMMON, "-x", "hip", "-shared", *MEMBENCH_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {MEMBENCH_SRC}:\n{r.stderr}")
    os.replace(tmp, MEMBENCH_LIB)
    if verbose:
        print("built", LIB, "and", MEMBENCH_LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
