"""The C++ natural class (include/nnh_natural.hpp) driven as the reference's
test.cpp drives natural: tests/cpp/natural_check prints one "key value" line
per result and every line is checked here against Python's integers, the
golden e digest and the golden factorial digest."""
import hashlib
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "natural_check")
GOLDEN = os.path.join(ROOT, "tests", "golden")

# operand shapes of natural_check.cpp, in its order
SHAPES = [(1, 1), (3, 2), (17, 5), (64, 64), (300, 40), (1000, 999), (4096, 1500), (20000, 7000), (70000, 30000)]


def test_natural_class_fails_loudly_without_device():
    """Without a GPU the class throws from its first device operation and the
    program exits non-zero naming the cause (no CPU fallback)."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present; the no-device path is checked on CPU hosts")
    assert os.path.exists(EXE), "tests/cpp/natural_check is not built (run __graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "libnnh:" in r.stderr and "device" in r.stderr


def splitmix64(state):
    M = (1 << 64) - 1
    while True:
        state = (state + 0x9E3779B97F4A7C15) & M
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        yield z ^ (z >> 31)


def random_int(gen, words):
    v = 0
    for i in range(words):
        v |= next(gen) << (64 * i)
    return v | (1 << (64 * words - 1))


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    assert os.path.exists(EXE), "tests/cpp/natural_check is not built (run __graft_entry__.build())"
    fact = str(tmp_path_factory.mktemp("natural") / "factorial.bin")
    print("natural_check: C++ class over libnnh, shapes", SHAPES, flush=True)
    r = subprocess.run([EXE, fact], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = dict(line.split(" ", 1) for line in r.stdout.splitlines())
    return out, fact


@pytest.mark.gpu
def test_pow_2_65536_decimal(run):
    out, _ = run
    s = out["pow2_65536"]
    assert len(s) == 19729 and s == str(2 ** 65536)


@pytest.mark.gpu
def test_operations_match_python_integers(run):
    out, _ = run
    gen = splitmix64(12345)
    for i, (an, bn) in enumerate(SHAPES):
        a, b = random_int(gen, an), random_int(gen, bn)
        h = lambda k: int(out[f"case{i}_{k}"], 16)
        assert h("a") == a and h("b") == b, i
        assert h("mul") == a * b, i
        assert h("sqr") == a * a, i
        assert h("div") == a // b, i
        assert h("mod") == a % b, i
        assert h("add") == a + b, i
        assert h("sub") == abs(a - b), i
        c = 1 if a > b else (-1 if a < b else 0)
        assert out[f"case{i}_cmp"] == f"{c} {-c} 0", i
        assert out[f"case{i}_identity"] == "1", i
        if an <= 1000:
            assert out[f"case{i}_dec"] == str(a * b), i
    x = random_int(gen, 9)
    m = random_int(gen, 3)
    assert int(out["compound_ops"], 16) == ((x * x) // 0x1234567) % m
    assert out["hex_upper"] == "0X" + format(0xABCDEF << 40, "X")


@pytest.mark.gpu
def test_e_recipe_matches_golden(run):
    out, _ = run
    g = json.load(open(os.path.join(GOLDEN, "e_digits.json")))["e_k14_D10000"]
    s = out["e_k14_D10000"]
    assert len(s) == g["length"] and s.startswith(g["head"])
    assert hashlib.sha256(s.encode()).hexdigest() == g["sha256"]


@pytest.mark.gpu
def test_factorial_matches_golden(run):
    out, fact = run
    g = json.load(open(os.path.join(GOLDEN, "factorial.json")))["1000000"]
    data = open(fact, "rb").read()
    assert int(out["factorial_1000000_words"]) == g["words"] == len(data) // 8
    assert hex(int.from_bytes(data[-8:], "little")) == g["top"]
    assert hashlib.sha256(data).hexdigest() == g["sha256"]


@pytest.mark.gpu
def test_reference_edge_cases(run):
    """1 - 2 = 1 (natural.h:324-340 subtracts the smaller from the larger);
    5 / 0 = 0, 0 / 0 = 1, 5 % 0 = 0 (natural.h:750-756)."""
    out, _ = run
    assert out["sub_smaller"] == "1"
    assert out["div_by_zero"] == "0 1 0"
