"""Builds libr2p.so (HIP kernels for gfx950 + C-ABI + C++ facade) in-tree.

Usage: python raw2png-conversionthrugpu_amd/build.py [--jobs N]
Each translation unit is compiled by hipcc --offload-arch=gfx950 into
_build/, then linked into raw2png-conversionthrugpu_amd/libr2p.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libr2p.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "r2p_kernels.hip",
    "r2p_capi.cpp",
    "r2p_strip.cpp",
    "r2p_batch.cpp",
    "facade.cpp",
    "pipeline.cpp",
]

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC] + COMMON + lang + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_facade_check()
    return LIB


def _build_facade_check() -> None:
    """tests/cpp/facade_check: C++ driver of the drop-in API (needs oracle/liboracle.so)."""
    src = os.path.join(REPO, "tests", "cpp", "facade_check.cpp")
    oracle_dir = os.path.join(REPO, "oracle")
    if not os.path.exists(src) or not os.path.exists(os.path.join(oracle_dir, "liboracle.so")):
        return
    out = os.path.join(REPO, "tests", "cpp", "facade_check")
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(REPO, "include"), "-I" + oracle_dir, src,
           "-L" + HERE, "-lr2p", "-L" + oracle_dir, "-loracle", "-lz",
           "-Wl,-rpath," + HERE, "-Wl,-rpath," + oracle_dir, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade_check build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    print(build(a.jobs))
    sys.exit(0)
