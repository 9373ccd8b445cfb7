"""Build the native library hilbert_amd/libv2rdm_hip.so with hipcc for gfx950.

Usage: python -m hilbert_amd.build   (or hilbert_amd.build.build()).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libv2rdm_hip.so")
SOURCES = ["layout.cpp", "constraints.cpp", "sell.cpp", "rdm.cpp", "projector.cpp", "solver.cpp", "sdp_host.cpp",
           "fcidump.cpp", "writers.cpp", "orbopt_driver.cpp", "abi.cpp",
           "kernels.hip",
           "eig_tri.hip", "refine.hip",
           "df.hip", "orbopt.hip", "noalloc.hip"]
ARCH = os.environ.get("V2H_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _flags(src: str):
    f = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"--offload-arch={ARCH}",
         "-I", os.path.join(HERE, "..", "include")]
    if src.endswith(".hip"):
        f += ["-x", "hip", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    return f


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *_flags(src), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
