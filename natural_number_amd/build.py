"""Build libnnh.so (HIP kernels + C ABI) for gfx950, in-tree.

Usage: python -m natural_number_amd.build   (or build() from Python)
The library lands next to this file so it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnnh.so")
ARCH = os.environ.get("NNH_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instruction selection stays off for every HIP translation unit.
HIP_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function"]

SOURCES = ["nnh_kernels.hip", "nnh_api.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


def build(verbose=False):
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = ["hipcc", "-c", src, "-o", obj] + COMMON
        if src.endswith(".hip"):
            cmd += HIP_FLAGS
        else:
            cmd += ["-x", "hip"] + HIP_FLAGS
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["hipcc", "-shared", "-o", tmp, f"--offload-arch={ARCH}"] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
