"""In-place products through the host-buffer calls: dst may overlap a or b
(a *= a, a *= b, natural.h:357-362), because both operands are staged into
device memory before any result word is written back.  Shapes cover the
schoolbook columns, the direct transform and the unbalanced split."""
import ctypes

import numpy as np
import pytest
from conftest import rand_limbs, to_int

pytestmark = pytest.mark.gpu

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _p(x, off=0):
    return ctypes.cast(x.ctypes.data + 8 * off, _u64p)


@pytest.mark.parametrize("an,bn", [(7, 3), (500, 450), (4096, 4096), (65536, 2000)])
def test_mul_in_place(nnh, an, bn):
    rng = np.random.default_rng(an * 31 + bn)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    want = to_int(a) * to_int(b)
    buf = np.zeros(an + bn, dtype=np.uint64)
    buf[:an] = a
    assert nnh.lib().nnh_mul(_p(buf), _p(buf), an, _p(b), bn) == 0   # dst aliases a
    assert to_int(buf) == want
    buf = np.zeros(an + bn, dtype=np.uint64)
    buf[:bn] = b
    assert nnh.lib().nnh_mul(_p(buf), _p(a), an, _p(buf), bn) == 0   # dst aliases b
    assert to_int(buf) == want


@pytest.mark.parametrize("an", [5, 384, 3000, 40000])
def test_sqr_and_self_product_in_place(nnh, an):
    rng = np.random.default_rng(an)
    a = rand_limbs(rng, an)
    want = to_int(a) ** 2
    buf = np.zeros(2 * an, dtype=np.uint64)
    buf[:an] = a
    assert nnh.lib().nnh_sqr(_p(buf), _p(buf), an) == 0
    assert to_int(buf) == want
    buf = np.zeros(2 * an, dtype=np.uint64)
    buf[:an] = a
    assert nnh.lib().nnh_mul(_p(buf), _p(buf), an, _p(buf), an) == 0  # a *= a
    assert to_int(buf) == want
