"""Pins of the CPU oracle (oracle/liboracle.so) — the restatement of the
reference CPU path — by checks that do not reuse its formulas: the
reference's own literals, closed forms, exact Python integers and brute-force
polynomial evaluation on tiny transforms."""
import ctypes
import random

import numpy as np
import pytest

from conftest import arr, from_int, rand_limbs, to_int

P = [27 << 56 | 1, 58 << 56 | 1, 87 << 56 | 1]
MASK = (1 << 64) - 1


def is_prime(n):
    if n < 2:
        return False
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def root_bases(oracle):
    rb = np.zeros(3 * 56, dtype=np.uint64)
    irb = np.zeros(3 * 56, dtype=np.uint64)
    oracle.L.nno_root_bases(arr(rb), arr(irb))
    return [[int(x) for x in rb[56 * i:56 * i + 56]] for i in range(3)], \
           [[int(x) for x in irb[56 * i:56 * i + 56]] for i in range(3)]


# ---- constants ---------------------------------------------------------------
def test_primes():
    for p, k in zip(P, [27, 58, 87]):
        assert p == k * 2**56 + 1 and is_prime(p)
    assert P[0] * P[1] * P[2] > 2**185


def test_root_bases_chain_and_reference_literals(oracle):
    rb, irb = root_bases(oracle)
    for i, p in enumerate(P):
        assert rb[i][0] == p - 1                                   # rb[0] = -1
        for j in range(55):
            assert rb[i][j + 1] ** 2 % p == rb[i][j]               # rb[j+1]^2 = rb[j]
        assert pow(rb[i][55], 2**55, p) == p - 1                   # order exactly 2^56
        for j in range(56):
            assert rb[i][j] * irb[i][j] % p == 1
    # literals of the reference's tables (ntt_info.h:83, 92, 101 and 113, 122, 131)
    assert [rb[i][1] for i in range(3)] == [0x0c55f5e087b66679, 0x0b5ea5c01435f877, 0x25eeaecc59203257]
    assert [irb[i][1] for i in range(3)] == [0x0eaa0a1f78499988, 0x2ea15a3febca078a, 0x31115133a6dfcdaa]
    # ntt_info.h:120 (p0 row, last entry) and ntt_info.h:129, 138
    assert irb[0][55] == 0x0c45d1745d1745d2
    assert irb[1][55] == 0x1d00000000000001
    assert irb[2][55] == 0x38f8915789157892


def test_derived_constants_match_reference_literals(oracle):
    c = np.zeros(16, dtype=np.uint64)
    oracle.L.nno_constants(arr(c))
    c = [int(x) for x in c]
    assert c[0:3] == P
    # mods_124 = ceil(2^124/p) (ntt_info.h:75-79)
    assert c[3:6] == [0x97b425ed097b425a, 0x469ee58469ee5846, 0x2f149902f149902f]
    # floor(2^(124+i)/p_i) (asmLinux.asm:926, 1075, 1236)
    assert c[6:9] == [0x97b425ed097b4259, 0x8d3dcb08d3dcb08a, 0xbc52640bc52640ba]
    # Garner constants of asmCRT (asmLinux.asm:1396-1425)
    assert c[9] == 0x1e73333333333335 and c[9] * P[0] % P[2] == 1
    assert c[10] == 0x599999999999999e
    assert c[11] == 0x772c234f72c234fa
    assert c[12] == 0x12b5ad6b5ad6b5aa and c[12] * P[0] * P[2] % P[1] == 1
    assert c[13] == 0x5294a5294a529495
    assert c[14] == 0x7200000000000001 and c[15] == 0x092d000000000000
    assert (c[15] << 64 | c[14]) == P[0] * P[2]


def test_twiddle_tables_exact(oracle):
    """root/iroot pairs (w_t, floor(w_t 2^64/p)+1) with w_t = prod rb[q+1]
    over set bits q of t (ntt_info.h:22-41), checked by exact arithmetic."""
    rb, irb = root_bases(oracle)
    s = 15
    for i, p in enumerate(P):
        for inv, base in ((0, rb[i]), (1, irb[i])):
            t = np.zeros(1 << (s - 1), dtype=np.uint64)
            oracle.L.nno_twiddles(arr(t), i, s, inv)
            for k in range(1 << (s - 2)):
                w = 1
                for q in range(k.bit_length()):
                    if k >> q & 1:
                        w = w * base[q + 1] % p
                assert int(t[2 * k]) == w
                assert int(t[2 * k + 1]) == (w << 64) // p + 1


# ---- stage outputs by brute force on tiny transforms ---------------------------
def _twid(base, p, g):
    w = 1
    for q in range(g.bit_length()):
        if g >> q & 1:
            w = w * base[q + 1] % p
    return w


@pytest.mark.parametrize("s,n", [(4, 5), (5, 16), (6, 40)])
def test_stages_brute_force(oracle, s, n):
    rb, irb = root_bases(oracle)
    rng = np.random.default_rng(s)
    a = rand_limbs(rng, n)
    b = rand_limbs(rng, (1 << s) - n)
    L = 1 << s
    fa = oracle.stage("forward", s, a)
    pw = oracle.stage("pointwise", s, a, b)
    iv = oracle.stage("inverse", s, a, b)
    for i, p in enumerate(P):
        av = [int(x) % p for x in a] + [0] * (L - n)
        bv = [int(x) % p for x in b] + [0] * (L - b.size)
        for g in range(L // 4):
            w = _twid(rb[i], p, g)
            for half, ww in ((0, w), (1, p - w)):        # X^2 - w_g, X^2 + w_g
                ra = [sum(av[j] * pow(ww, j // 2, p) for j in range(e, L, 2)) % p for e in (0, 1)]
                rbv = [sum(bv[j] * pow(ww, j // 2, p) for j in range(e, L, 2)) % p for e in (0, 1)]
                got = [int(fa[i][4 * g + 2 * half + e]) % p for e in (0, 1)]
                assert got == ra, (i, g, half)
                prod = [(ra[0] * rbv[0] + ww * ra[1] * rbv[1]) % p, (ra[0] * rbv[1] + ra[1] * rbv[0]) % p]
                assert [int(pw[i][4 * g + 2 * half + e]) for e in (0, 1)] == prod
        conv = [0] * L
        for x in range(L):
            for y in range(L):
                conv[(x + y) % L] += av[x] * bv[y]
        assert [int(v) for v in iv[i]] == [c % p for c in conv]


# ---- products ----------------------------------------------------------------
PRODUCT_SIZES = [(1, 1), (2, 1), (5, 3), (31, 31), (32, 32), (47, 47), (48, 48), (33, 200), (383, 383),
                 (384, 384), (384, 383), (1024, 1024), (1500, 400), (4096, 4097), (10000, 700),
                 (65536, 65536), (100000, 1000), (200000, 3000)]


@pytest.mark.parametrize("an,bn", PRODUCT_SIZES)
def test_mul_exact(oracle, an, bn):
    rng = np.random.default_rng(an * 31 + bn)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    assert to_int(oracle.mul(a, b)) == to_int(a) * to_int(b)


@pytest.mark.parametrize("n", [1, 2, 47, 48, 383, 384, 3000, 50000])
def test_sqr_exact(oracle, n):
    a = rand_limbs(np.random.default_rng(n), n)
    assert to_int(oracle.sqr(a)) == to_int(a) ** 2


def test_mul_arms_agree(oracle):
    rng = np.random.default_rng(9)
    for an, bn in [(400, 400), (2000, 500), (3000, 3000)]:
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        base = np.zeros(an + bn, dtype=np.uint64)
        ntt = np.zeros(an + bn, dtype=np.uint64)
        oracle.L.nno_mul_base(arr(base), arr(a), an, arr(b), bn)
        oracle.L.nno_mul_ntt(arr(ntt), arr(a), an, arr(b), bn)
        assert np.array_equal(base, ntt)


def test_l3_blocked_scale_mod_checks(oracle):
    """s = 21 (> NTT_scale_L3_threshold, ntt_info.h:179): residues mod random primes."""
    rng = np.random.default_rng(21)
    a, b = rand_limbs(rng, (1 << 19) + 5), rand_limbs(rng, (1 << 19) + 3)
    d = oracle.mul(a, b)
    A, B, D = to_int(a), to_int(b), to_int(d)
    for q in [2**61 - 1, 2**89 - 1, 1000000007, (1 << 64) - 59]:
        assert D % q == (A % q) * (B % q) % q
    assert D.bit_length() <= 64 * (a.size + b.size)


def test_closed_forms(oracle):
    n = 5000
    ones = np.full(n, MASK, dtype=np.uint64)
    sq = oracle.sqr(ones)
    exp = np.zeros(2 * n, dtype=np.uint64)
    exp[0] = 1
    exp[n] = MASK - 1
    exp[n + 1:] = MASK
    assert np.array_equal(sq, exp)
    m = 3001
    got = to_int(oracle.mul(ones, ones[:m].copy()))
    assert got == ((1 << 64 * n) - 1) * ((1 << 64 * m) - 1)
    x = np.zeros(700, dtype=np.uint64)
    x[-1] = 1 << 63
    y = np.zeros(900, dtype=np.uint64)
    y[-1] = 1 << 5
    assert to_int(oracle.mul(x, y)) == 1 << (64 * 699 + 63 + 64 * 899 + 5)


@pytest.mark.parametrize("s,an,bn", [(10, 800, 700), (12, 4096, 4096), (13, 5000, 8192), (16, 40000, 60000)])
def test_cyclic_mode(oracle, s, an, bn):
    """save() wraps when the product exceeds the transform (ntt_info.h:238-242):
    the result is a*b mod (B^L - 1)."""
    rng = np.random.default_rng(s * an)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    L = 1 << s
    got = oracle.ntt_product(a, b, s)
    assert got.size == L
    m = (1 << (64 * L)) - 1
    assert to_int(got) % m == to_int(a) * to_int(b) % m


def test_exact_mode_through_one_transform(oracle):
    rng = np.random.default_rng(3)
    a, b = rand_limbs(rng, 700), rand_limbs(rng, 300)
    got = oracle.ntt_product(a, b, 10)
    assert to_int(got) == to_int(a) * to_int(b) and got.size == 1000


# ---- division, decimal output, e ------------------------------------------------
@pytest.mark.parametrize("an,bn", [(1, 1), (5, 1), (3, 2), (10, 4), (300, 200), (2000, 700), (6000, 2500),
                                   (9000, 9000), (20000, 3000)])
def test_divmod_exact(oracle, an, bn):
    rng = np.random.default_rng(an + 7 * bn)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    q, r = oracle.divmod(a, b)
    A, B = to_int(a), to_int(b)
    assert to_int(q) == A // B and to_int(r) == A % B


def test_divmod_reference_defect(oracle):
    """Divisors whose normalised leading words are all ones (here B^n - 1)
    make the reference's reciprocal (natural.h:620-687) shrink its accuracy
    to 0 and memset a wrapped negative size: undefined in the reference, so
    the oracle reports it instead of faulting."""
    from conftest import ReferenceDefect

    ones = np.full(3000, MASK, dtype=np.uint64)
    b = np.full(1200, MASK, dtype=np.uint64)
    with pytest.raises(ReferenceDefect):
        oracle.divmod(ones, b)


def test_divmod_adversarial(oracle):
    b = np.full(1200, MASK, dtype=np.uint64)
    b[-1] = 1 << 62      # leading word not all ones after normalisation
    ones = np.full(3000, MASK, dtype=np.uint64)
    q, r = oracle.divmod(ones, b)
    assert to_int(q) == to_int(ones) // to_int(b) and to_int(r) == to_int(ones) % to_int(b)
    p2 = np.zeros(2000, dtype=np.uint64)
    p2[-1] = 1
    b2 = np.zeros(600, dtype=np.uint64)
    b2[-1] = 1 << 40
    q, r = oracle.divmod(p2, b2)
    assert to_int(q) == to_int(p2) // to_int(b2) and to_int(r) == 0


@pytest.mark.parametrize("x", [0, 7, 10**18 - 1, 10**18, 10**19 + 12345, 2**64 - 1, 2**64, 10**36, 3**3000])
def test_to_dec_small(oracle, x):
    assert oracle.to_dec(from_int(x, max(1, (x.bit_length() + 63) // 64))) == str(x)


def test_to_dec_large(oracle):
    rng = random.Random(4)
    for bits in [5000, 70000, 400000]:
        x = rng.getrandbits(bits) | (1 << (bits - 1))
        assert oracle.to_dec(from_int(x, (bits + 63) // 64)) == str(x)


def test_pow_2_65536_known_answer(oracle):
    """The reference demo prints pow(natural(2), 65536) (test.cpp:9)."""
    out = np.zeros(1026, dtype=np.uint64)
    n = oracle.L.nno_pow(arr(out), out.size, 2, 65536)
    v = to_int(out[:n])
    assert v == 2**65536
    s = oracle.to_dec(out[:n])
    assert len(s) == 19729
    assert s.startswith("2003529930406846464979072351560255750447825475569751419265016973")
    assert s == str(2**65536)


def test_e_digits_exact(oracle):
    """test.cpp:12-19 recipe at k=10, 2000 digits, against the exact rational
    sum_{j<2^k} 1/j! and the known prefix of e."""
    k, D = 10, 2000
    got = oracle.e_digits(k, D)
    M = 1 << k
    den = 1
    for j in range(1, M):
        den *= j
    num, t = 0, den
    for j in range(0, M):
        if j:
            t //= j
        num += t
    assert got == str(num * 10**D // den)
    assert got.startswith("27182818284590452353602874713526624977572470936999595749669676277240766303535")


# ---- pi (natural.h:859-902, test.cpp:22-32) -------------------------------------
def _words(x):
    return max(1, (x.bit_length() + 63) // 64)


def _py_calc_pi(b, e):
    """Exact-integer restatement of calc_pi (natural.h:874-902)."""
    if e - b == 1:
        den = b * b * b * 0x26DD041D878000
        numc = (2 * b - 1) * (6 * b - 1) * (6 * b - 5)
        num = (0x207E2DA6 * b + 0xCF6371) * numc
        return num, den, numc
    m = (b + e) >> 1
    num, den, numc = _py_calc_pi(b, m)
    num1, den1, numc1 = _py_calc_pi(m, e)
    if e - b == 2:
        return abs(numc * num1 - num * den1), den * den1, numc * numc1
    return num * den1 + numc * num1, den * den1, numc * numc1


def _py_pi_digits(k, digits, acc):
    num, den, numc = _py_calc_pi(1, (1 << k) | 1)
    sn, sd = 4001, 40                       # sqrt_10005 (natural.h:859-872)
    while _words(sn) + _words(sd) - 2 < acc:
        sn, sd = sd * sd * 10005 + sn * sn, (sn * sd) << 1
    num = abs(0xCF6371 * den - num)
    den = den * (0x68380 * sn * 10**digits)
    num = num * sd
    return str(den // num)


def _oracle_pi(oracle, k, digits, acc):
    buf = ctypes.create_string_buffer(digits + 64)
    oracle.L.nno_pi_digits.restype = ctypes.c_uint64
    oracle.L.nno_pi_digits.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.c_uint64]
    n = oracle.L.nno_pi_digits(buf, digits + 64, k, digits, acc)
    return buf.raw[:n].decode()


@pytest.mark.parametrize("k,digits,acc", [(4, 100, 8), (6, 500, 30), (8, 2000, 120), (10, 6000, 350)])
def test_pi_recipe_exact(oracle, k, digits, acc):
    assert _oracle_pi(oracle, k, digits, acc) == _py_pi_digits(k, digits, acc)


def test_pi_digits_vs_mpmath(oracle):
    """The recipe's digits are pi's (up to the last few, which the truncated
    series and square root may not fix): independent pin by mpmath."""
    import mpmath
    D = 20000
    got = _oracle_pi(oracle, 11, D, 1200)
    mpmath.mp.dps = D + 30
    ref = mpmath.nstr(mpmath.pi, D + 20, strip_zeros=False).replace(".", "")
    assert got.startswith("31415926535897932384626433832795028841971693993751")
    assert got[:D - 5] == ref[:D - 5]
