"""Build the gfx950 shared library in-tree, lbm_amd/liblbm_hip.so, and the C++
command-line driver linked against it, lbm_amd/lbm_cli.

    python -m lbm_amd.build

hipcc compiles for gfx950 only.  Every HIP source is built with
-ffp-contract=off (the step must round like the reference's f32 arithmetic: no
fused multiply-add) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liblbm_hip.so")
SOURCES = ["lbm_grid.hip", "lbm_batch.hip", "unet.hip", "lbm_f64.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


CLI = os.path.join(HERE, "lbm_cli")
CLI_SOURCE = os.path.join(CSRC, "lbm_cli.cpp")


def build_cli(verbose: bool = False) -> str:
    """The C++ command-line driver (host code only, linked against the library)."""
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    tmp = CLI + ".tmp"
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), CLI_SOURCE,
           "-L", HERE, "-l:liblbm_hip.so", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("building lbm_cli failed")
    os.replace(tmp, CLI)
    return CLI


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
