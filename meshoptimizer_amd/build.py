"""Build recipe for libmeshopt_hip.so (gfx950).

hipcc compiles the HIP kernels and the host orchestrator into one shared library next to this
file.  Floating-point flags are part of the contract: no FP contraction (the reference's x86
build forms no FMA, and a fused multiply-add changes the collapse order), IEEE division and
square root, denormals kept, and no packed-FP32 instruction selection.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import concurrent.futures as futures

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmeshopt_hip.so")
BUILD_DIR = os.path.join(HERE, "_build")

SOURCES = ["prims.hip", "setup.hip", "pass.hip", "prune.hip", "batch.hip", "batch_narrow.hip", "batch_mid.hip", "batch_wide.hip", "sloppy.hip", "solve.hip", "clod.hip", "host.cpp"]

ARCH = os.environ.get("MESHOPT_HIP_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [hipcc(), *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link libmeshopt_hip.so; returns the library path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    with futures.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
