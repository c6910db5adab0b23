"""The oracle built under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/cpp/oracle_san.cpp): products at every threshold (transform against
schoolbook), squares, shifts, additions, the Newton division (q b + r = a),
pow and decimal output (2^65536, the reference's demo), calc_e (the hex
digits of e), all without a memory or undefined-behaviour error."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_san")
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-march=x86-64-v3", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "tests", "cpp", "oracle_san.cpp"), os.path.join(ROOT, "oracle", "nn_oracle.cpp"),
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "oracle_san ok" in r.stdout
