"""The modular arithmetic as the kernels compile it (asm Montgomery products,
the fused-addend lazy forward butterfly, twiddle-1 butterflies, the asm base
multiply, Garner's reconstruction), run on the GPU over random and boundary
inputs and checked on the host against exact 128-bit arithmetic
(tests/cpp/arith_dev_check.hip)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "arith_dev_check")


@pytest.mark.gpu
def test_device_arithmetic_exact():
    assert os.path.exists(EXE), "tests/cpp/arith_dev_check is not built (run __graft_entry__.build())"
    print("device arithmetic check: 3 primes x 2^18 cases", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
