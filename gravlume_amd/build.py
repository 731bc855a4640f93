"""Builds libgravlume_hip.so (gfx950) in-tree with hipcc.

    python gravlume_amd/build.py            # compile
    python gravlume_amd/build.py --asm      # also print per-kernel register usage
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgravlume_hip.so")
SOURCES = ["gl_trace.hip", "gl_accel.hip", "gl_scene.cpp", "gl_spectral.cpp", "gl_flops.cpp",
           "gl_vjp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for every HIP translation unit.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=on", "-Wno-unused-result",
]


HOST_CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-c"]


def build(save_asm: bool = False, verbose: bool = False) -> str:
    """Host-only translation units (.cpp) go through the host compiler, the
    HIP units (.hip) through hipcc; hipcc links the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not src.endswith(".cpp") or not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src[:-4] + ".o")
        res = subprocess.run([HOST_CXX, *HOST_FLAGS, "-o", obj, path], capture_output=True,
                             text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"host compile of {src} failed")
        objs.append(obj)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not src.endswith(".hip") or not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src[:-4] + ".o")
        cmd = [HIPCC, *COMMON, "-c", "-o", obj, path]
        if save_asm:
            cmd += ["-Rpass-analysis=kernel-resource-usage"]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose and res.stderr:
            sys.stderr.write(res.stderr)
        objs.append(obj)
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("linking libgravlume_hip.so failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(save_asm="--asm" in sys.argv, verbose=True)
