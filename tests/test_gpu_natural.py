"""Parity of the device-resident natural layer (nnh_divmod, nnh_pow,
nnh_to_dec, nnh_calc_e, nnh_e_digits) with Python integers, the CPU oracle
and the committed e digests.  Integer work: bit-exact."""
import hashlib
import json
import os
import random

import numpy as np
import pytest
from conftest import ReferenceDefect, from_int, rand_limbs, to_int

pytestmark = pytest.mark.gpu
MASK = (1 << 64) - 1
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "e_digits.json")))


def _norm(x):
    x = np.asarray(x, dtype=np.uint64)
    n = x.size
    while n > 1 and x[n - 1] == 0:
        n -= 1
    return x[:n]


# ---- division (natural.h:566-769) -----------------------------------------------
@pytest.mark.parametrize("an,bn", [(1, 1), (5, 1), (3, 2), (10, 4), (300, 200), (2000, 700), (6000, 2500),
                                   (9000, 9000), (20000, 3000), (40000, 39000)])
def test_divmod_exact(nnh, an, bn):
    rng = np.random.default_rng(an + 7 * bn)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    q, r = nnh.divmod_(a, b)
    A, B = to_int(a), to_int(b)
    assert to_int(q) == A // B and to_int(r) == A % B
    assert q[-1] != 0 or q.size == 1
    assert r[-1] != 0 or r.size == 1


@pytest.mark.parametrize("an,bn", [(70000, 20000), (600000, 250000), (1 << 20, 300000)])
def test_divmod_large_vs_oracle(nnh, oracle, an, bn):
    """Sizes where the reciprocal's products wrap (cyclic save) and
    div_iterative folds modulo B^L - 1 (natural.h:717-724)."""
    rng = np.random.default_rng(an ^ bn)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    q, r = nnh.divmod_(a, b)
    oq, orr = oracle.divmod(a, b)
    assert np.array_equal(q, _norm(oq)) and np.array_equal(r, _norm(orr))


def test_divmod_edge_cases(nnh):
    a = from_int(12345, 1)
    zero = np.zeros(1, dtype=np.uint64)
    q, r = nnh.divmod_(a, zero)          # natural.h:750-755
    assert to_int(q) == 0 and to_int(r) == 0
    q, r = nnh.divmod_(zero, zero)
    assert to_int(q) == 1 and to_int(r) == 0
    big = rand_limbs(np.random.default_rng(3), 50)
    q, r = nnh.divmod_(a, big)           # a < b
    assert to_int(q) == 0 and to_int(r) == 12345
    q, r = nnh.divmod_(big, big)
    assert to_int(q) == 1 and to_int(r) == 0


def test_divmod_adversarial(nnh):
    b = np.full(1200, MASK, dtype=np.uint64)
    b[-1] = 1 << 62
    ones = np.full(3000, MASK, dtype=np.uint64)
    q, r = nnh.divmod_(ones, b)
    assert to_int(q) == to_int(ones) // to_int(b) and to_int(r) == to_int(ones) % to_int(b)
    p2 = np.zeros(2000, dtype=np.uint64)
    p2[-1] = 1
    b2 = np.zeros(600, dtype=np.uint64)
    b2[-1] = 1 << 40
    q, r = nnh.divmod_(p2, b2)
    assert to_int(q) == to_int(p2) // to_int(b2) and to_int(r) == 0


def test_divmod_reference_defect_reported(nnh, oracle):
    """B^n - 1 divisors: the reference's reciprocal is undefined
    (natural.h:675-677); the oracle and the GPU layer both report it."""
    ones = np.full(3000, MASK, dtype=np.uint64)
    b = np.full(1200, MASK, dtype=np.uint64)
    with pytest.raises(ReferenceDefect):
        oracle.divmod(ones, b)
    with pytest.raises(nnh.NNHError, match="status 5"):
        nnh.divmod_(ones, b)


# ---- pow and decimal output ------------------------------------------------------
def test_pow_known_answers(nnh):
    assert to_int(nnh.pow_(from_int(2, 1), 65536)) == 2**65536       # test.cpp:9
    assert to_int(nnh.pow_(from_int(10, 1), 100000)) == 10**100000
    assert to_int(nnh.pow_(from_int(3**40, 2), 777)) == (3**40) ** 777
    assert to_int(nnh.pow_(from_int(7, 1), 0)) == 1


@pytest.mark.parametrize("x", [0, 7, 10**18 - 1, 10**18, 10**19 + 12345, 2**64 - 1, 2**64, 10**36, 10**36 - 1,
                               10**288 - 1, 10**288, 3**3000, 10**576 + 1])
def test_to_dec_small(nnh, x):
    assert nnh.to_dec(from_int(x, max(1, (x.bit_length() + 63) // 64))) == str(x)


def test_pow_2_65536_printed(nnh):
    s = nnh.to_dec(nnh.pow_(from_int(2, 1), 65536))
    assert len(s) == 19729 and s == str(2**65536)
    assert s.startswith("2003529930406846464979072351560255750447825475569751419265016973")


@pytest.mark.parametrize("bits", [5000, 70000, 400000])
def test_to_dec_random(nnh, bits):
    rng = random.Random(bits)
    x = rng.getrandbits(bits) | (1 << (bits - 1))
    assert nnh.to_dec(from_int(x, (bits + 63) // 64)) == str(x)


@pytest.mark.parametrize("words", [1 << 17, 3 << 18])
def test_to_dec_large_vs_oracle(nnh, oracle, words):
    a = rand_limbs(np.random.default_rng(words), words)
    assert nnh.to_dec(a) == oracle.to_dec(a)


def test_to_dec_powers_of_ten(nnh):
    """Runs of zeros and a leading 1 stress the correction loop."""
    for e in [17, 18, 19, 35, 36, 37, 575, 576, 577, 9215, 9216, 9217, 150000]:
        for d in (-1, 0, 1):
            x = 10**e + d
            assert nnh.to_dec(from_int(x, (x.bit_length() + 63) // 64)) == str(x)


def test_to_dec_largest_pieces_per_level(nnh):
    """Pieces at the top of each level's range: B_(i+1) - 1 = B_i^2 - 1 and
    q B_i +- 1, where B_i = 10^(18 2^i) and every base below the top one
    uses a reciprocal derived from the next base's."""
    for i in range(5, 12):
        b = 10 ** (18 << i)
        for x in (b * b - 1, 999999937 * b - 1, 999999937 * b + 1, (b - 1) * b):
            assert nnh.to_dec(from_int(x, (x.bit_length() + 63) // 64)) == str(x), i


# ---- e (natural.h:841-857, test.cpp:12-19)----------------------------------------
@pytest.mark.parametrize("k", [6, 10, 14, 17])
def test_calc_e_matches_oracle(nnh, oracle, k):
    num, den = nnh.calc_e(k)
    cap = (24 << k) // 64 + 64
    on = np.zeros(cap, dtype=np.uint64)
    od = np.zeros(cap, dtype=np.uint64)
    import ctypes
    nn_, dn_ = ctypes.c_uint64(), ctypes.c_uint64()
    u64p = ctypes.POINTER(ctypes.c_uint64)
    oracle.L.nno_calc_e(on.ctypes.data_as(u64p), cap, od.ctypes.data_as(u64p), cap, k, ctypes.byref(nn_),
                        ctypes.byref(dn_))
    assert np.array_equal(num, on[:nn_.value]) and np.array_equal(den, od[:dn_.value])


def test_calc_e_small_exact(nnh):
    k = 7
    num, den = nnh.calc_e(k)
    M = 1 << k
    D = 1
    for j in range(1, M):
        D *= j
    N, t = 0, D
    for j in range(M):
        if j:
            t //= j
        N += t
    assert to_int(den) == D and to_int(num) == N


@pytest.mark.parametrize("name", ["e_k14_D10000", "e_k17_D100000", "e_k18_D1000000", "e_k21_D10000000",
                                  "e_k24_D100000000"])
def test_e_digits_golden(nnh, name):
    g = GOLDEN[name]
    s, ms = nnh.e_digits(g["k"], g["digits"])
    assert len(s) == g["length"]
    assert s[:60] == g["head"] and s[-60:] == g["tail"]
    assert hashlib.sha256(s.encode()).hexdigest() == g["sha256"]


def test_e_digits_vs_oracle(nnh, oracle):
    s, _ = nnh.e_digits(12, 3000)
    assert s == oracle.e_digits(12, 3000)
    assert s.startswith("271828182845904523536028747135266249775724709369995")


# ---- factorial (natural.h:795-839) ---------------------------------------------
@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 10, 31, 100, 1000, 20000, 200000])
def test_factorial_exact(nnh, n):
    import math
    assert to_int(nnh.factorial(n)) == math.factorial(n)


@pytest.mark.parametrize("n", ["1000000", "10000000"])
def test_factorial_golden(nnh, n):
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "factorial.json")))[n]
    f = nnh.factorial(int(n))
    assert f.size == g["words"] and hex(int(f[-1])) == g["top"]
    assert hashlib.sha256(f.astype("<u8").tobytes()).hexdigest() == g["sha256"]


# ---- pi (natural.h:859-902, test.cpp:22-32) --------------------------------------
@pytest.mark.parametrize("k", [2, 4, 8, 12, 16])
def test_calc_pi_matches_oracle(nnh, oracle, k):
    import ctypes
    num, den, numc = nnh.calc_pi(k)
    cap = (340 << k) // 64 + 64
    bufs = [np.zeros(cap, dtype=np.uint64) for _ in range(3)]
    sizes = np.zeros(3, dtype=np.uint64)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    oracle.L.nno_calc_pi.argtypes = [u64p, u64p, u64p, ctypes.c_uint64, ctypes.c_int, u64p]
    oracle.L.nno_calc_pi(*(b.ctypes.data_as(u64p) for b in bufs), cap, k, sizes.ctypes.data_as(u64p))
    for got, b, n in zip((num, den, numc), bufs, sizes):
        assert np.array_equal(got, b[:int(n)])


def test_pi_digits_small_exact(nnh):
    from test_oracle import _py_pi_digits
    for k, D, acc in [(4, 100, 8), (6, 500, 30), (8, 2000, 120)]:
        s, _ = nnh.pi_digits(k, D, acc)
        assert s == _py_pi_digits(k, D, acc)


@pytest.mark.parametrize("name", ["pi_k14_D200000_acc12000", "pi_k17_D1000000_acc60000",
                                  "pi_k20_D10000000_acc600000"])
def test_pi_digits_golden(nnh, name):
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "pi_digits.json")))[name]
    s, _ = nnh.pi_digits(g["k"], g["digits"], g["acc"])
    assert len(s) == g["length"] and s[:60] == g["head"] and s[-60:] == g["tail"]
    assert hashlib.sha256(s.encode()).hexdigest() == g["sha256"]


def test_shutdown_releases_and_recovers(nnh, oracle):
    """nnh_shutdown frees the twiddle tables, workspaces and the natural
    layer's buffer pool; the next calls rebuild them and give the same
    results."""
    a = rand_limbs(np.random.default_rng(77), 5000)
    before = (nnh.e_digits(12, 3000)[0], nnh.to_dec(a), to_int(nnh.mul(a, a)))
    assert nnh.lib().nnh_shutdown() == 0
    after = (nnh.e_digits(12, 3000)[0], nnh.to_dec(a), to_int(nnh.mul(a, a)))
    assert after == before
    assert after[0] == oracle.e_digits(12, 3000) and after[1] == oracle.to_dec(a)
    assert after[2] == to_int(a) ** 2


@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 4097, 65537, 1000000])
def test_divmod_one_word_divisor_vs_python(nnh, n):
    """One-word divisors (div_base, natural.h:566-574): the chunked
    remainder scan above 2048 words against Python's integers, divisors with
    and without the top bit, 1 and 2^64 - 1."""
    rng = np.random.default_rng(n)
    a = rand_limbs(rng, n)
    A = to_int(a)
    for d in (1, 3, 10**18, (1 << 63) + 12345, (1 << 64) - 1, int(rng.integers(2, 2**62))):
        q, r = nnh.divmod_(a, np.array([d], dtype=np.uint64))
        qq, rr = divmod(A, d)
        assert to_int(q) == qq and to_int(r) == rr, (n, d)
