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
