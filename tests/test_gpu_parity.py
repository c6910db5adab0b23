"""Parity of the HIP product path (libnnh.so through its C ABI) with the CPU
oracle and the committed golden digests.  Integer work: bit-exact."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import from_int, rand_limbs, to_int

pytestmark = pytest.mark.gpu

P = [27 << 56 | 1, 58 << 56 | 1, 87 << 56 | 1]
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "products.json")))


def say(*a):
    print(*a, flush=True)


def digest(w):
    return hashlib.sha256(np.ascontiguousarray(w, dtype="<u8").tobytes()).hexdigest()


SIZES = [
    (1, 1), (2, 1), (3, 2), (31, 31), (32, 32), (100, 3), (383, 383), (383, 5000),   # schoolbook arm
    (384, 384), (385, 400), (1024, 1024), (1500, 400), (2048, 2048), (2049, 2048),  # s = 10 .. 13
    (5000, 4000), (20000, 300), (40000, 33000), (65536, 65536), (100000, 1000),     # split and top passes
    (300000, 5000), (1 << 19, 1 << 17),
]


@pytest.mark.parametrize("an,bn", SIZES)
def test_mul_matches_oracle(nnh, oracle, an, bn):
    rng = np.random.default_rng(an * 7919 + bn)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    say(f"mul {an}x{bn}")
    assert np.array_equal(nnh.mul(a, b), oracle.mul(a, b))


@pytest.mark.parametrize("n", [1, 47, 383, 384, 1000, 4096, 5000, 1 << 17])
def test_sqr_matches_oracle(nnh, oracle, n):
    rng = np.random.default_rng(n)
    a = rand_limbs(rng, n)
    say(f"sqr {n}")
    assert np.array_equal(nnh.sqr(a), oracle.sqr(a))


def test_random_sizes(nnh, oracle):
    rng = np.random.default_rng(2024)
    for it in range(40):
        an = int(2 ** rng.uniform(0, 17))
        bn = int(2 ** rng.uniform(0, 17))
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        say(f"random {it}: {an}x{bn}")
        assert np.array_equal(nnh.mul(a, b), oracle.mul(a, b)), (an, bn)


def test_adversarial_operands(nnh, oracle):
    ones = np.full(5000, 2**64 - 1, dtype=np.uint64)
    single = np.zeros(3000, dtype=np.uint64)
    single[-1] = 1 << 63
    top = np.zeros(4000, dtype=np.uint64)
    top[-1] = 1
    for a, b in [(ones, ones[:4096]), (single, ones), (top, top), (single, single)]:
        say(f"adversarial {a.size}x{b.size}")
        assert np.array_equal(nnh.mul(a, b), oracle.mul(a, b))


def test_config1_exact(nnh, oracle):
    g = GOLDEN["c1_2^20x2^20"]
    a = nnh.splitmix64(g["a"]["seed"], g["a"]["words"])
    b = nnh.splitmix64(g["b"]["seed"], g["b"]["words"])
    say("config1 2^20 x 2^20 bits")
    d = nnh.mul(a, b)
    assert np.array_equal(d, oracle.mul(a, b))
    assert digest(d) == g["sha256"]


@pytest.mark.parametrize("name", ["c2_2^30x2^30", "c2_2^30x2^22", "c2_2^28x2^28", "mid_2^25x2^23"])
def test_golden_full_size(nnh, name):
    g = GOLDEN[name]
    a = nnh.splitmix64(g["a"]["seed"], g["a"]["words"])
    b = nnh.splitmix64(g["b"]["seed"], g["b"]["words"])
    say(f"golden {name}")
    d = nnh.mul(a, b)
    for i, v in g["sample"].items():
        assert hex(int(d[int(i)])) == v, i
    assert digest(d) == g["sha256"]


def test_all_ones_square_closed_form(nnh):
    # (B^n - 1)^2 = B^2n - 2 B^n + 1: words [1, 0 x (n-1), 2^64-2, (2^64-1) x (n-1)]
    n = 1 << 22
    a = np.full(n, 2**64 - 1, dtype=np.uint64)
    say("all-ones 2^28-bit square (via mul and sqr)")
    exp = np.empty(2 * n, dtype=np.uint64)
    exp[0] = 1
    exp[1:n] = 0
    exp[n] = 2**64 - 2
    exp[n + 1:] = 2**64 - 1
    assert np.array_equal(nnh.mul(a, a.copy()), exp)
    assert np.array_equal(nnh.sqr(a), exp)


@pytest.mark.parametrize("n", [1000, 40000, (1 << 20) + 3, (1 << 21) - 1])
def test_all_ones_square_short_crt_tiles(nnh, n):
    # Outputs below 2^22 words reconstruct in 512-word CRT tiles; the upper
    # half of (B^n - 1)^2 is all ones, so a carry entering it crosses every
    # tile there.  Square via mul and sqr.
    a = np.full(n, 2**64 - 1, dtype=np.uint64)
    say(f"all-ones {n}-word square")
    exp = np.empty(2 * n, dtype=np.uint64)
    exp[0] = 1
    exp[1:n] = 0
    exp[n] = 2**64 - 2
    exp[n + 1:] = 2**64 - 1
    assert np.array_equal(nnh.mul(a, a.copy()), exp)
    assert np.array_equal(nnh.sqr(a), exp)


@pytest.mark.parametrize("n,m", [(300000, 70000), (70001, 1 << 19)])
def test_all_ones_unbalanced_closed_form(nnh, n, m):
    # (B^n - 1)(B^m - 1) = B^(n+m) - B^n - B^m + 1
    a = np.full(n, 2**64 - 1, dtype=np.uint64)
    b = np.full(m, 2**64 - 1, dtype=np.uint64)
    say(f"all-ones {n}x{m}")
    B = 64
    want = (1 << (B * (n + m))) - (1 << (B * n)) - (1 << (B * m)) + 1
    assert to_int(nnh.mul(a, b)) == want


@pytest.mark.parametrize("scale,n", [(8, 100), (10, 600), (12, 4096), (13, 5000), (16, 40000), (21, 1 << 20)])
def test_forward_transform_stage(nnh, oracle, scale, n):
    """Device transform == alpha * reference transform (mod p), every word."""
    rng = np.random.default_rng(scale)
    a = rand_limbs(rng, n)
    say(f"forward stage s={scale} n={n}")
    dev = nnh.transform_forward(a, scale).reshape(3, 1 << scale)
    ref = oracle.stage("forward", scale, a)
    al = nnh.alpha(scale)
    for i in range(3):
        p = P[i]
        d = dev[i].astype(object) % p
        r = ref[i].astype(object) % p
        assert np.all(d == (r * al[i]) % p), i
        # lazy ranges: [0, 4p) for p0 and p1 (4p < 2^64), [0, 2p) for p2
        bound = 4 * p if i < 2 else 2 * p
        assert np.all(dev[i].astype(object) < bound), i


@pytest.mark.parametrize("scale,an,bn", [(10, 800, 700), (12, 3000, 2500), (14, 16384, 9000), (16, 65536, 65536)])
def test_cyclic_matches_oracle(nnh, oracle, scale, an, bn):
    rng = np.random.default_rng(scale + an)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    say(f"cyclic s={scale} {an}x{bn}")
    got = nnh.mul_cyclic(a, b, scale)
    exp = oracle.ntt_product(a, b, scale)
    assert np.array_equal(got, exp)
    # and the value is a*b mod (B^L - 1)
    L = 1 << scale
    m = (1 << (64 * L)) - 1
    assert to_int(got) % m == (to_int(a) * to_int(b)) % m


def test_cyclic_all_ones(nnh, oracle):
    a = np.full(1 << 12, 2**64 - 1, dtype=np.uint64)
    say("cyclic all-ones s=12")
    assert np.array_equal(nnh.mul_cyclic(a, a.copy(), 12), oracle.ntt_product(a, a.copy(), 12))


@pytest.mark.parametrize("scale,an,bn", [(10, 500, 400), (14, 9000, 7000), (20, 600000, 400000)])
def test_transform_reuse(nnh, scale, an, bn):
    rng = np.random.default_rng(an)
    a = rand_limbs(rng, an)
    fa = nnh.transform_forward(a, scale)
    for k in range(2):
        b = rand_limbs(rng, bn - k * 7)
        say(f"reuse s={scale} {an}x{b.size}")
        got = nnh.mul_transformed(fa, an, b, scale)
        assert to_int(got) == to_int(a) * to_int(b)
