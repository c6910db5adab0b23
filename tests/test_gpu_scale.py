"""Products at the largest transform lengths (2^26, 2^28 and 2^30 words per
prime: top passes of 8 + 8 levels and, at 2^30, three passes), beyond the
sizes the oracle checks in seconds.  Size-independent properties pin them:
the residues mod 2^64 - 1 and mod 2^64 + 1 of a product equal the products of
the operands' residues (a wrong twiddle, a lost carry or a misplaced pass
breaks them with overwhelming probability), and the low 64 words are exact
(they depend only on the operands' low words)."""
import numpy as np
import pytest

from conftest import to_int

pytestmark = pytest.mark.gpu

M1, P1 = (1 << 64) - 1, (1 << 64) + 1


def say(*a):
    print(*a, flush=True)


def halves_sum(x):
    """sum of x (uint64 words) as an exact Python int."""
    lo = int(np.sum(x & np.uint64(0xFFFFFFFF), dtype=np.uint64))
    hi = int(np.sum(x >> np.uint64(32), dtype=np.uint64))
    return lo + (hi << 32)


def res_m1(x):  # B = 2^64 = 1 mod 2^64 - 1
    return halves_sum(x) % M1


def res_p1(x):  # B = -1 mod 2^64 + 1
    return (halves_sum(x[0::2]) - halves_sum(x[1::2])) % P1


@pytest.mark.parametrize("scale", [26, 28, 30])
def test_large_product_residues(nnh, scale):
    n = 1 << (scale - 1)
    rng = np.random.default_rng(scale)
    a = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    b = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    a[-1] |= np.uint64(1)
    b[-1] |= np.uint64(1)
    say(f"scale {scale}: {n} x {n} words, plan {nnh.transform_plan(scale)}")
    p = nnh.mul(a, b)
    assert p.size == 2 * n
    assert res_m1(p) == res_m1(a) * res_m1(b) % M1
    assert res_p1(p) == res_p1(a) * res_p1(b) % P1
    k = 64
    assert to_int(p[:k]) == to_int(a[:k]) * to_int(b[:k]) % (1 << (64 * k))
    del p
