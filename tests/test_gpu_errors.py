"""Argument errors at the C ABI: each call returns its status (no exception
crosses the boundary, nothing is enqueued for an invalid request), names
the cause in nnh_last_error(), and leaves the library working: a product
after the errors is exact."""
import ctypes

import numpy as np
import pytest
from conftest import rand_limbs, to_int

pytestmark = pytest.mark.gpu

ARG, DOMAIN = 1, 5
_u64p = ctypes.POINTER(ctypes.c_uint64)


def _p(x):
    return x.ctypes.data_as(_u64p)


def _err(nnh):
    return nnh.lib().nnh_last_error().decode()


def test_argument_errors_then_exact_product(nnh):
    L = nnh.lib()
    rng = np.random.default_rng(5)
    a, b = rand_limbs(rng, 5000), rand_limbs(rng, 700)
    out = np.zeros(a.size + b.size, dtype=np.uint64)
    n = ctypes.c_uint64()
    cases = [
        ("mul null dst", lambda: L.nnh_mul(None, _p(a), a.size, _p(b), b.size), ARG, "null pointer"),
        ("mul empty operand", lambda: L.nnh_mul(_p(out), _p(a), 0, _p(b), b.size), ARG, "empty operand"),
        ("sqr null operand", lambda: L.nnh_sqr(_p(out), None, a.size), ARG, "null pointer"),
        ("unknown arm", lambda: L.nnh_mul_algo(_p(out), _p(a), a.size, _p(b), b.size, 99, 0), ARG, "unknown algorithm"),
        # a piece transform of 2^9 = 512 words cannot hold the 700-word operand
        ("chunk piece scale too small", lambda: L.nnh_mul_algo(_p(out), _p(a), a.size, _p(b), b.size, 5, 9), ARG,
         "bad piece scale"),
        ("cyclic scale below range", lambda: L.nnh_mul_cyclic(_p(out), _p(a), a.size, _p(b), b.size, 2), ARG,
         "bad arguments"),
        ("cyclic operand longer than the transform",
         lambda: L.nnh_mul_cyclic(_p(out), _p(a), a.size, _p(b), b.size, 12), ARG, "longer than the transform"),
        ("pow output too small", lambda: L.nnh_pow(_p(out), 2, ctypes.byref(n), _p(b), b.size, 3), ARG, "too small"),
        ("divmod empty divisor", lambda: L.nnh_divmod(_p(out), 10, ctypes.byref(n), _p(out), 10, ctypes.byref(n),
                                                      _p(a), a.size, _p(b), 0), ARG, "empty operand"),
    ]
    for name, call, code, text in cases:
        rc = call()
        assert rc == code, (name, rc, _err(nnh))
        assert text in _err(nnh), (name, _err(nnh))
    # decimal output into a buffer too small for the digits
    buf = ctypes.create_string_buffer(16)
    assert L.nnh_to_dec(buf, 16, ctypes.byref(n), _p(b), b.size) == ARG and "too small" in _err(nnh)
    # a batch with one bad request is refused as a whole
    d = nnh.DeviceWords(64)
    reqs = (nnh.Req * 2)(nnh.Req(d.ptr, d.ptr, 8, d.ptr, 8), nnh.Req(None, d.ptr, 8, d.ptr, 8))
    h = ctypes.c_void_p()
    assert L.nnh_batch_create(reqs, 2, 0, ctypes.byref(h)) == ARG and "batch request 1" in _err(nnh)
    d.free()
    # the library still computes
    assert to_int(nnh.mul(a, b)) == to_int(a) * to_int(b)
    assert to_int(nnh.mul_algo(a, b, nnh.ALGO_NTT_CHUNK, 11)) == to_int(a) * to_int(b)
