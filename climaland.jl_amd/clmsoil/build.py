"""Build libclmsoil.so (gfx950) in-tree with hipcc.

Usage: python -m clmsoil.build   (from the climaland.jl_amd directory), or
call build_library() from Python.  The library lands next to the package:
climaland.jl_amd/libclmsoil.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_ROOT = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
LIB_PATH = os.path.join(PKG_ROOT, "libclmsoil.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are disabled for all HIP code in this project
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libclmsoil.so)")


def build_library(out: str = LIB_PATH, extra: list[str] | None = None, verbose: bool = False) -> str:
    src = os.path.join(CSRC, "clm_soil.hip")
    tmp = out + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *(extra or []), "-I", os.path.join(REPO_ROOT, "include"),
           "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-6000:]}")
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    odir = os.path.join(REPO_ROOT, "oracle")
    res = subprocess.run(["make", "-s", "-C", odir], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr[-4000:]}")
    return os.path.join(odir, "liborc.so")


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
