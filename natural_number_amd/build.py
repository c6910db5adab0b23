"""Build libnnh.so (HIP kernels + C ABI) for gfx950, in-tree.

Usage: python -m natural_number_amd.build   (or build() from Python)
The library lands next to this file so it travels with the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnnh.so")
ARCH = os.environ.get("NNH_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instruction selection stays off for every HIP translation unit.
HIP_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function"]

SOURCES = ["nnh_mid.hip", "nnh_mid_lazy.hip", "nnh_top.hip", "nnh_top_lazy.hip", "nnh_crt.hip", "nnh_mlp.hip", "nnh_peak.hip",
           "nnh_nat.hip", "nnh_api.cpp", "nnh_natural.cpp", "nnh_plan.cpp", "nnh_hostio.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


def _compile(src, verbose):
    obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
    cmd = ["hipcc", "-c", src, "-o", obj] + COMMON
    if not src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += HIP_FLAGS
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    return obj


def build(verbose=False):
    """Compile every translation unit in parallel and link libnnh.so."""
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    _run(["hipcc", "-shared", "-o", tmp, f"--offload-arch={ARCH}"] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_arith_check(verbose=False):
    """The device arithmetic check run by tests/test_gpu_arith.py (test
    infrastructure, built beside its source)."""
    src_dir = os.path.join(os.path.dirname(HERE), "tests", "cpp")
    exe = os.path.join(src_dir, "arith_dev_check")
    cmd = ["hipcc", "arith_dev_check.hip", "-o", exe + ".tmp", "-O3", "-std=c++20", f"--offload-arch={ARCH}"] + HIP_FLAGS
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=src_dir, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    os.replace(exe + ".tmp", exe)
    return exe


def build_natural_check(verbose=False):
    """The C++ class check run by tests/test_gpu_natural_cpp.py (host code
    over include/nnh_natural.hpp, linked to the in-tree libnnh.so)."""
    root = os.path.dirname(HERE)
    src_dir = os.path.join(root, "tests", "cpp")
    exe = os.path.join(src_dir, "natural_check")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "natural_check.cpp", "-o", exe + ".tmp",
           "-I", os.path.join(root, "include"), "-L", HERE, "-lnnh", "-Wl,-rpath,$ORIGIN/../../natural_number_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=src_dir, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    os.replace(exe + ".tmp", exe)
    return exe


def extract_bindings(md_path, out_path):
    """Every cpp code block of INTEGRATION.md's binding section, concatenated
    into one include file (the bindings a maintainer would add to the
    reference's natural.h)."""
    text = open(md_path).read()
    head = "## The binding a maintainer would add to natural.h"
    start = text.index(head) + len(head)
    nxt = text.find("\n## ", start)
    sec = text[start:nxt if nxt >= 0 else len(text)]
    blocks = []
    pos = 0
    while True:
        i = sec.find("```cpp\n", pos)
        if i < 0:
            break
        j = sec.index("\n```", i + 7)
        blocks.append(sec[i + 7:j + 1])
        pos = j + 4
    if not blocks:
        raise RuntimeError("no cpp blocks in the binding section of " + md_path)
    body = "// Generated from INTEGRATION.md by natural_number_amd/build.py; do not edit.\n" + "\n".join(blocks)
    tmp = out_path + ".tmp"
    with open(tmp, "w") as f:
        f.write(body)
    os.replace(tmp, out_path)
    return len(blocks)


def build_binding_check(verbose=False, out_dir=None):
    """The INTEGRATION.md bindings compiled against a stand-in natural whose
    CPU bodies are the oracle (tests/cpp/binding_check.cpp; test
    infrastructure, run by tests/test_bindings.py)."""
    root = os.path.dirname(HERE)
    src_dir = os.path.join(root, "tests", "cpp")
    out_dir = out_dir or src_dir
    inc = os.path.join(out_dir, "binding_snippets.inc")
    extract_bindings(os.path.join(root, "INTEGRATION.md"), inc)
    exe = os.path.join(out_dir, "binding_check")
    cmd = ["g++", "-O2", "-std=c++20", "-march=x86-64-v3", "-Wall", "-Wextra", "-Wno-unused-function",
           os.path.join(src_dir, "binding_check.cpp"), os.path.join(root, "oracle", "nn_oracle.cpp"),
           "-o", exe + ".tmp", "-I", out_dir, "-I", os.path.join(root, "include"), "-I", os.path.join(root, "oracle"),
           "-L", HERE, "-lnnh", "-Wl,-rpath," + HERE]
    if out_dir == src_dir:
        cmd[-1] = "-Wl,-rpath,$ORIGIN/../../natural_number_amd"
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=src_dir, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
