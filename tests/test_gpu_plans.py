"""Level splits of the transform (the chunk size and the top-pass split)
beyond the built-in formula: single products and batches under other valid
splits against the oracle, the autotuner's timing and installation, and the
plan table's file form."""
import numpy as np
import pytest

from conftest import rand_limbs
from test_gpu_batch import DevBatch

pytestmark = pytest.mark.gpu

# (scale, chunk bits, passes): extremes of the candidate space
SPLITS = [(13, 10, [1, 2]), (13, 11, [1, 1]), (16, 10, [3, 3]), (16, 12, [3, 1]), (16, 11, [4, 1]),
          (20, 10, [3, 3, 4]), (20, 11, [7, 1, 1]), (21, 12, [3, 3, 3]), (22, 10, [4, 4, 4]), (22, 12, [8, 2])]


def say(*a):
    print(*a, flush=True)


@pytest.fixture
def at(nnh):
    from natural_number_amd import autotune as at_

    saved = {}

    def set_(s, batched, cb, passes):
        saved.setdefault((s, batched), at_.plan(s, batched))
        at_.set_plan(s, batched, cb, passes)

    yield at_, set_
    for (s, batched), (cb, passes, tuned) in saved.items():
        at_.set_plan(s, batched, cb if tuned else 0, passes if tuned else [])


@pytest.mark.parametrize("s,cb,passes", SPLITS)
def test_products_under_other_splits(nnh, oracle, at, s, cb, passes):
    at_, set_ = at
    rng = np.random.default_rng(s * 100 + cb)
    set_(s, False, cb, passes)
    L = 1 << s
    for an, bn in [(L // 2, L // 2), (L // 2 - 7, L // 3), ((L * 3) // 4, L // 4 + 1)]:
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        got = nnh.mul_algo(a, b, nnh.ALGO_NTT)
        assert np.array_equal(got, oracle.mul(a, b)), (s, cb, passes, an, bn)
    a = rand_limbs(rng, L // 2)
    assert np.array_equal(nnh.sqr(a), oracle.sqr(a))


@pytest.mark.parametrize("s,cb,passes", [x for x in SPLITS if x[0] <= 20])
def test_batches_under_other_splits(nnh, oracle, at, s, cb, passes):
    at_, set_ = at
    rng = np.random.default_rng(s * 7 + cb)
    set_(s, True, cb, passes)
    L = 1 << s
    ops = [(rand_limbs(rng, int(L // 2 - rng.integers(0, L // 8))), rand_limbs(rng, int(L // 2 - rng.integers(0, L // 8))))
           for _ in range(6)]
    shared = rand_limbs(rng, L // 4)
    ops += [(rand_limbs(rng, L // 2), shared) for _ in range(3)]
    db = DevBatch(nnh, ops)
    for graph in (False, True):
        b = nnh.Batch(db.reqs, graph=graph, reference=True)
        db.clear()
        b.run()
        nnh.synchronize()
        for (x, y), got in zip(ops, db.results()):
            assert np.array_equal(got, oracle.mul(x, y)), (s, cb, passes, x.size, y.size, graph)
        b.close()


def test_tuner_times_and_installs(nnh, at):
    at_, set_ = at
    for s, batched in ((14, False), (16, True)):
        set_(s, batched, 0, [])
        r = at_.tune(s, batched, reps=3)
        say(f"s={s} batched={batched}: " + ", ".join(f"{[x['chunk_bits']] + x['passes']} {x['ms']:.4f}" for x in r))
        assert len(r) >= 3 and all(x["ms"] > 0 for x in r)
        best = min(r, key=lambda x: x["ms"])
        cb, passes, tuned = at_.plan(s, batched)
        if best["ms"] < 0.99 * r[0]["ms"]:  # installed only when it beats the formula by more than 1 %
            assert tuned and [cb] + passes == [best["chunk_bits"]] + best["passes"]
        else:
            assert not tuned


def test_plan_file_round_trip(nnh, at, tmp_path):
    at_, set_ = at
    set_(18, False, 11, [4, 3])
    set_(17, True, 10, [4, 3])
    p = tmp_path / "plans.txt"
    at_.save(p)
    set_(18, False, 0, [])
    set_(17, True, 0, [])
    at_.load(p)
    assert at_.plan(18, False) == (11, [4, 3], True)
    assert at_.plan(17, True) == (10, [4, 3], True)
    head, *rest = p.read_text().splitlines()
    (tmp_path / "other.txt").write_text("nnh-plans 1 gfx000 1\n" + "\n".join(rest) + "\n")
    with pytest.raises(nnh.NNHError, match="another device"):
        at_.load(tmp_path / "other.txt")
    with pytest.raises(nnh.NNHError):
        at_.set_plan(18, False, 8, [5, 5])  # chunks below 2^10 words beside top passes are refused
