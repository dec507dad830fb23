"""Build libsemx.so (HIP kernels for gfx950 + host setup) in-tree.

Usage: python -m semtex_amd.build   (or semtex_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsemx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["semx_setup.cpp", "semx_io.cpp", "semx_solve.hip", "semx_dns.hip"]

# Packed FP32 instructions are switched off for all device code.
DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose: bool = True) -> str:
    """Compile every source (in parallel, one compiler process each) and
    link libsemx.so."""
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(bdir, src + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC, "-c", path, "-o", obj] + COMMON + DEVICE_FLAGS
        else:
            cmd = ["g++", "-c", path, "-o", obj, "-pthread"] + COMMON
        print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, "compile " + " ".join(failed))
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-o", tmp] + objs + DEVICE_FLAGS + ["-pthread"])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
