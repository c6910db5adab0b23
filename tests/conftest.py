import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

sys.set_int_max_str_digits(0)

_u64p = ctypes.POINTER(ctypes.c_uint64)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def arr(x):
    return x.ctypes.data_as(_u64p)


class ReferenceDefect(Exception):
    pass


class Oracle:
    """ctypes view of oracle/liboracle.so (the CPU restatement; checker only)."""

    def __init__(self):
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        u64, i32 = ctypes.c_uint64, ctypes.c_int
        for name, args, res in [
            ("nno_mul", [_u64p, _u64p, u64, _u64p, u64], u64),
            ("nno_sqr", [_u64p, _u64p, u64], u64),
            ("nno_mul_base", [_u64p, _u64p, u64, _u64p, u64], u64),
            ("nno_mul_ntt", [_u64p, _u64p, u64, _u64p, u64], u64),
            ("nno_sqr_ntt", [_u64p, _u64p, u64], u64),
            ("nno_ntt_product", [_u64p, u64, _u64p, u64, _u64p, u64, i32], u64),
            ("nno_stage_load", [_u64p, _u64p, u64, i32], None),
            ("nno_stage_forward", [_u64p, _u64p, u64, i32], None),
            ("nno_stage_pointwise", [_u64p, _u64p, u64, _u64p, u64, i32], None),
            ("nno_stage_inverse", [_u64p, _u64p, u64, _u64p, u64, i32], None),
            ("nno_twiddles", [_u64p, i32, i32, i32], None),
            ("nno_root_bases", [_u64p, _u64p], None),
            ("nno_constants", [_u64p], None),
            ("nno_divmod", [_u64p, _u64p, _u64p, u64, _u64p, u64, _u64p, _u64p], i32),
            ("nno_pow", [_u64p, u64, u64, u64], u64),
            ("nno_to_dec", [ctypes.c_char_p, u64, _u64p, u64], u64),
            ("nno_e_digits", [ctypes.c_char_p, u64, i32, u64], u64),
            ("nno_calc_e", [_u64p, u64, _u64p, u64, i32, _u64p, _u64p], None),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        self.L = L

    def mul(self, a, b):
        out = np.zeros(a.size + b.size, dtype=np.uint64)
        self.L.nno_mul(arr(out), arr(a), a.size, arr(b), b.size)
        return out

    def sqr(self, a):
        out = np.zeros(2 * a.size, dtype=np.uint64)
        self.L.nno_sqr(arr(out), arr(a), a.size)
        return out

    def ntt_product(self, a, b, scale):
        cap = max(1 << scale, a.size + b.size)
        out = np.zeros(cap, dtype=np.uint64)
        n = self.L.nno_ntt_product(arr(out), cap, arr(a), a.size, arr(b), b.size, scale)
        return out[:n]

    def stage(self, name, scale, a, b=None):
        out = np.zeros(3 << scale, dtype=np.uint64)
        f = getattr(self.L, "nno_stage_" + name)
        if b is None:
            f(arr(out), arr(a), a.size, scale)
        else:
            f(arr(out), arr(a), a.size, arr(b), b.size, scale)
        return out.reshape(3, 1 << scale)

    def divmod(self, a, b):
        q = np.zeros(a.size, dtype=np.uint64)
        r = np.zeros(b.size, dtype=np.uint64)
        qn, rn = ctypes.c_uint64(0), ctypes.c_uint64(0)
        st = self.L.nno_divmod(arr(q), arr(r), arr(a), a.size, arr(b), b.size, ctypes.byref(qn), ctypes.byref(rn))
        if st == 2:
            raise ReferenceDefect("reference division undefined for this divisor")
        return q[: qn.value], r[: rn.value]

    def to_dec(self, a):
        n = self.L.nno_to_dec(None, 0, arr(a), a.size)
        buf = ctypes.create_string_buffer(n + 1)
        self.L.nno_to_dec(buf, n + 1, arr(a), a.size)
        return buf.value.decode()

    def e_digits(self, k, digits):
        n = digits + 64
        buf = ctypes.create_string_buffer(n + 1)
        m = self.L.nno_e_digits(buf, n + 1, k, digits)
        return buf.value.decode()[:m]


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


@pytest.fixture(scope="session")
def nnh():
    import natural_number_amd as nn

    nn.lib()
    return nn


def to_int(w):
    return int.from_bytes(np.ascontiguousarray(w, dtype=np.uint64).tobytes(), "little")


def from_int(x, n):
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype=np.uint64).copy()


def rand_limbs(rng, n, top_nonzero=True):
    a = rng.integers(0, 2**64, size=n, dtype=np.uint64, endpoint=False)
    if top_nonzero and a[-1] == 0:
        a[-1] = 1
    return a
