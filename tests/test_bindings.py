"""The reference-side bindings INTEGRATION.md shows (natural::_mul, _sqr,
_div and operator<< over the C ABI), compiled from that file against a
stand-in natural with basic_natural's resize/move semantics whose CPU
bodies are the oracle (tests/cpp/binding_check.cpp), and run three ways:
without a device (every call falls back to the CPU bodies), on the GPU, and
on the GPU with a failure injected into each operation's device call
(nnh_debug_fail_next), where the fallback must give identical results.
Cases: a *= a, a * b and a *= b across the 384-word threshold, sqr,
a %= b (remainder aliasing the dividend, natural.h:783-787), a / b, a % b,
a /= b, a /= a, a %= a, a one-word divisor, decimal output, the
transform-reuse helpers over device transform handles (mul_save, mul_load
including a cyclic product, sqr_save) and calc_e written as the reference
writes it over those helpers, each compared with the oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "binding_check")


def _run(exe, mode):
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    return r


def test_bindings_compile_and_fall_back_without_device(tmp_path):
    """Extract the snippets from INTEGRATION.md afresh, compile them, and run
    them where nnh_init fails (no device): every operation takes the
    reference body and matches the oracle."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present; the no-device path is checked on CPU hosts")
    from natural_number_amd import build as nb

    exe = nb.build_binding_check(out_dir=str(tmp_path))
    r = _run(exe, "nogpu")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "binding_check ok (nogpu)" in r.stdout


def test_snippets_extracted_from_integration_doc(tmp_path):
    from natural_number_amd import build as nb

    out = tmp_path / "s.inc"
    n = nb.extract_bindings(os.path.join(ROOT, "INTEGRATION.md"), str(out))
    text = out.read_text()
    assert n >= 2
    for fn in ("natural& natural::_mul(", "natural& natural::_sqr(", "natural& natural::_div(",
               "operator<<(std::ostream& os, const natural& n)"):
        assert fn in text


@pytest.mark.gpu
def test_bindings_on_gpu():
    assert os.path.exists(EXE), "tests/cpp/binding_check is not built (run __graft_entry__.build())"
    r = _run(EXE, "gpu")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "binding_check ok (gpu)" in r.stdout


@pytest.mark.gpu
def test_bindings_fall_back_on_injected_failure():
    assert os.path.exists(EXE), "tests/cpp/binding_check is not built (run __graft_entry__.build())"
    r = _run(EXE, "inject")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "binding_check ok (inject)" in r.stdout
