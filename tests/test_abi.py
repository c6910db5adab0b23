"""The C ABI library loads and exports every entry point include/nnh.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nnh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nnh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["nnh_init", "nnh_mul", "nnh_sqr", "nnh_mul_device_async", "nnh_mul_cyclic",
              "nnh_transform_forward", "nnh_mul_transformed", "nnh_threshold_words"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    import natural_number_amd as nn

    lib = ctypes.CDLL(nn.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_threshold_is_reference_value():
    import natural_number_amd as nn

    assert nn.threshold_words() == 384  # NTTmul_threshold, ntt_info.h:6


def test_missing_library_fails_loudly(monkeypatch):
    import natural_number_amd as nn

    monkeypatch.setattr(nn, "_lib", None)
    monkeypatch.setattr(nn, "LIB_PATH", "/nonexistent/libnnh.so")
    try:
        nn.lib()
    except nn.NNHError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("loading a missing library must raise")


def test_transform_plan_splits_every_level():
    """Every transform level runs exactly once: chunk_bits - 1 in the fused
    chunk kernel, the rest in top passes of at most 8 levels (the base
    multiply takes the last level); the 2^25-word product uses 2^10 chunks
    and passes of 8 and 7 levels."""
    import natural_number_amd as nn

    assert nn.transform_plan(25) == (10, [8, 7])
    for s in range(2, 31):
        cb, passes = nn.transform_plan(s)
        assert cb + sum(passes) == s
        assert all(1 <= p <= 8 for p in passes)
        assert cb <= 12 and (cb == s or cb >= 10)
        assert len(passes) == (0 if s <= 12 else (s - 12 + 7) // 8)


def test_cpp_class_header_compiles(tmp_path):
    """include/nnh_natural.hpp compiles on its own (C++17 and C++20, warnings
    as errors); the GPU behaviour is covered by test_gpu_natural_cpp.py."""
    import subprocess

    src = tmp_path / "use.cpp"
    src.write_text('#include "nnh_natural.hpp"\n'
                   "int main() { nnh::natural a(3), b = a; return cmp(a, b) + (int)(a + b).size(); }\n")
    for std in ("c++17", "c++20"):
        r = subprocess.run(["g++", f"-std={std}", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                            "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_transform_plan_formula():
    """The built-in level split (no device, no tuned table): up to 2^20 words
    the largest chunk (2^12) and balanced passes deeper first; above, the
    fewest passes of at most 8 levels and a chunk of at least 2^10 words."""
    import ctypes

    import natural_number_amd as nn

    L = nn.lib()
    cb, np_, lev = ctypes.c_int(), ctypes.c_int(), (ctypes.c_int * 4)()
    for s, want in [(8, (8, [])), (12, (12, [])), (13, (12, [1])), (20, (12, [8])), (21, (10, [6, 5])),
                    (25, (10, [8, 7])), (28, (12, [8, 8])), (30, (10, [7, 7, 6]))]:
        assert L.nnh_transform_plan(s, ctypes.byref(cb), ctypes.byref(np_), lev) == 0
        assert (cb.value, list(lev)[:np_.value]) == want, s
