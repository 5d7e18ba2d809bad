"""Build the native library in-tree: ``chuanflow_amd/_lib/libchuanflow_amd.so``.

HIP sources are compiled for gfx950 with hipcc; host-only C++ with g++.
Every object keeps ``-ffp-contract=off`` so no fused multiply-add changes a
rounding the reference does not make (the assembly, geometry and flux
kernels are bit-exact against the reference build).  Packed FP32 ops are
disabled for all device code.

Usage: ``python -m chuanflow_amd.build`` (or ``build_all()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB_NAME = "libchuanflow_amd.so"
ARCH = os.environ.get("CFX_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["dev/krylov.hip", "dev/pcg.hip", "dev/mg.hip", "dev/dic.hip", "dev/mpnn.hip", "dev/train.hip", "dev/fvops.hip", "dev/mesh.hip", "dev/simple.hip", "dev/noalloc.hip", "cfx_api.hip"]
CXX_SOURCES = ["cf/PolyMesh.cpp", "cf/DetailBlock.cpp", "cf/MeshIO.cpp", "cfx_host_api.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-I" + CSRC, "-I" + os.path.join(REPO, "include")]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _obj(src):
    return os.path.join(OBJ_DIR, src.replace("/", "_").rsplit(".", 1)[0] + ".o")


def build_all(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([hipcc, "-c", os.path.join(CSRC, s), "-o", _obj(s)] + COMMON + HIP_FLAGS)
    for s in CXX_SOURCES:
        jobs.append(["g++", "-c", os.path.join(CSRC, s), "-o", _obj(s)] + COMMON)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr, file=sys.stderr)
    out = os.path.join(OUT_DIR, LIB_NAME)
    tmp = out + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp]
         + [_obj(s) for s in HIP_SOURCES + CXX_SOURCES])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_all(verbose=True))
