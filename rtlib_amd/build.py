"""Builds librtpg.so (host C++ + gfx950 HIP kernels) in-tree with hipcc.

Usage: python -m rtlib_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "librtpg.so"

SOURCES = ["bvh.cpp", "api.cpp", "sdtree_host.cpp", "hashgrid_host.cpp", "mlp_host.cpp", "render.hip", "sdtree.hip",
           "hashgrid.hip", "mlp.hip", "trace_noalloc.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + str(PKG.parent / "include"),
]
# Device flags: gfx950 only; hardware float atomics; packed-FP32 instructions
# disabled (they have produced silent wrong results on these boxes).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str) -> Path:
    s = CSRC / src
    o = BUILD / (src + ".o")
    cmd = [HIPCC] + COMMON
    if src == "bvh.cpp":
        cmd += ["-x", "c++"]
    else:
        # Host files include the HIP runtime headers; compiling them as HIP
        # keeps one toolchain configuration for the whole library.
        cmd += HIP_FLAGS + ["-x", "hip"]
    cmd += ["-c", str(s), "-o", str(o)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return o


def build(jobs: int = 8, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-o", str(tmp)] + [str(o) for o in objs] + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args(argv)
    build(a.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
