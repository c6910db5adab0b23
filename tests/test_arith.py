"""The device modular arithmetic (nnh_arith.hpp: Montgomery reduction for
p = k*2^56+1, butterflies, Garner) compiled for the host and checked against
exact 128-bit arithmetic."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_device_arithmetic_on_host(tmp_path):
    exe = tmp_path / "arith_check"
    src = os.path.join(ROOT, "tests", "cpp", "arith_check.cpp")
    cmd = ["g++", "-std=c++20", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++", src,
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
