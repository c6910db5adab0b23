"""Host-buffer calls at sizes that go through the pinned staging ring
(transfers of 8 MiB and more, in 16 MiB chunks split over copy threads):
operand and result sizes that are no multiple of a chunk, of the copy
threads' share or of a word block; dst aliasing an operand (a *= a, which
the ring must not break: both operands are read before any result word is
written); squares; and a device round trip.  Products are pinned by the
size-independent checks of the large-product test (residues mod 2^64 - 1
and 2^64 + 1, exact low words) and against the device-pointer path, whose
parity with the oracle the other GPU tests establish."""
import ctypes

import numpy as np
import pytest

from test_gpu_scale import M1, P1, res_m1, res_p1

pytestmark = pytest.mark.gpu

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _p(x):
    return x.ctypes.data_as(_u64p)


def _low_words(a, b, k):
    """exact low k words of a * b from the operands' low words"""
    x = int.from_bytes(a[:k].tobytes(), "little") * int.from_bytes(b[:k].tobytes(), "little")
    return np.frombuffer((x % (1 << (64 * k))).to_bytes(8 * k, "little"), dtype=np.uint64)


def _device_product(nnh, a, b):
    da, db, do = nnh.DeviceWords(a.size, a), nnh.DeviceWords(b.size, b), nnh.DeviceWords(a.size + b.size)
    nnh.mul_device_async(do.ptr, da.ptr, a.size, db.ptr, b.size)
    nnh.synchronize()
    return do.download()


@pytest.mark.parametrize("an,bn", [((1 << 21) + 12345, (1 << 20) + 777), ((1 << 22) + 3, (1 << 22) - 5)])
def test_staged_product(nnh, an, bn):
    rng = np.random.default_rng(an ^ bn)
    a = rng.integers(0, 2**64, size=an, dtype=np.uint64)
    b = rng.integers(0, 2**64, size=bn, dtype=np.uint64)
    a[-1] |= np.uint64(1)
    b[-1] |= np.uint64(1)
    out = np.zeros(an + bn, dtype=np.uint64)
    assert nnh.lib().nnh_mul(_p(out), _p(a), an, _p(b), bn) == 0, nnh.lib().nnh_last_error()
    assert res_m1(out) == res_m1(a) * res_m1(b) % M1
    assert res_p1(out) == res_p1(a) * res_p1(b) % P1
    assert np.array_equal(out[:64], _low_words(a, b, 64))
    assert np.array_equal(out, _device_product(nnh, a, b))


def test_staged_in_place_and_square(nnh):
    n = (1 << 21) + 999
    rng = np.random.default_rng(7)
    a = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    a[-1] |= np.uint64(1)
    want = _device_product(nnh, a, a)
    buf = np.zeros(2 * n, dtype=np.uint64)
    buf[:n] = a
    assert nnh.lib().nnh_mul(_p(buf), _p(buf), n, _p(buf), n) == 0   # a *= a: dst aliases both operands
    assert np.array_equal(buf, want)
    buf = np.zeros(2 * n, dtype=np.uint64)
    buf[:n] = a
    assert nnh.lib().nnh_sqr(_p(buf), _p(buf), n) == 0
    assert np.array_equal(buf, want)


def test_staged_round_trip_after_shutdown(nnh):
    """The ring is released by nnh_shutdown and rebuilt by the next staged call."""
    n = (1 << 21) + 1
    rng = np.random.default_rng(11)
    a = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    b = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    out1 = np.zeros(2 * n, dtype=np.uint64)
    assert nnh.lib().nnh_mul(_p(out1), _p(a), n, _p(b), n) == 0
    assert nnh.lib().nnh_shutdown() == 0
    out2 = np.zeros(2 * n, dtype=np.uint64)
    assert nnh.lib().nnh_mul(_p(out2), _p(a), n, _p(b), n) == 0
    assert np.array_equal(out1, out2)
    assert np.array_equal(out1, _device_product(nnh, a, b))
