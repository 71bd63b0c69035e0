"""Build libbntt.so (the HIP kernels and the C ABI of include/bntt.h) in-tree,
and libbntt_calltime.so beside it (the reference's timing protocol in C over
the host-array entry points, used by bench.py).

    python benchntt_amd/build.py

Every HIP source is compiled for gfx950 only, with packed-FP32 code generation
disabled, and linked into benchntt_amd/libbntt.so with hidden visibility except
for the bntt_* entry points.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbntt.so")
SOURCES = ["api.hip", "blas.hip", "ntt.hip", "batch.hip", "serve.hip", "io.hip", "probe.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
CALLTIME_LIB = os.path.join(HERE, "libbntt_calltime.so")

CFLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-fvisibility=hidden",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_calltime(verbose)
    return LIB


def build_calltime(verbose: bool = False) -> str:
    """libbntt_calltime.so: the reference's per-call timing protocol in C over
    the host-array entry points (csrc/call_timing.c), linked to libbntt.so."""
    tmp = CALLTIME_LIB + ".tmp"
    cmd = [CC, "-O2", "-std=gnu11", "-Wall", "-Werror", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "call_timing.c"), "-L", HERE, "-lbntt", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{CC} failed on call_timing.c:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CALLTIME_LIB)
    if verbose:
        print(f"built {CALLTIME_LIB}")
    return CALLTIME_LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
