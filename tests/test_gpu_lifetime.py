"""Device memory shared between calls: batches keep the twiddle tables their
launches (and captured graphs) read when a larger product grows the
device's tables or nnh_shutdown releases them, and device-pointer products
queued on different caller streams take turns in the shared workspace.
Every product is compared with the oracle."""
import numpy as np
import pytest

from conftest import rand_limbs
from test_gpu_batch import DevBatch, make_ops

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def check_batch(db, ops, oracle, what):
    for i, ((a, b), got) in enumerate(zip(ops, db.results())):
        exp = oracle.sqr(a) if b is a else oracle.mul(a, b)
        assert np.array_equal(got, exp), (what, i, a.size, b.size)


def grow_twiddles(nnh, words):
    """One product whose transform (2^(bit_length(2 words - 2))) is larger than any before."""
    a = nnh.DeviceWords(words, nnh.splitmix64(91, words))
    out = nnh.DeviceWords(2 * words)
    nnh.mul_device_async(out.ptr, a.ptr, words, a.ptr, words)
    nnh.synchronize()
    for x in (a, out):
        x.free()


def test_graph_batch_keeps_its_twiddles(nnh, oracle):
    """A graph batch planned right after nnh_shutdown (tables for 2^16 words),
    then a 2^25-word product (new, larger tables), then replays, then
    nnh_shutdown again and another replay: every replay is exact."""
    rng = np.random.default_rng(2024)
    nnh._check(nnh.lib().nnh_shutdown())
    sizes = [(int(2 ** rng.uniform(0, 14.9)), int(2 ** rng.uniform(0, 14.9))) for _ in range(40)]
    sizes += [(384, 384), (20000, 12000), (30000, 30000)]
    ops = make_ops(rng, sizes, square_every=7)
    db = DevBatch(nnh, ops)
    say("graph batch after shutdown")
    b = nnh.Batch(db.reqs, graph=True)
    b.run()
    nnh.synchronize()
    check_batch(db, ops, oracle, "first replay")
    say("2^25-word product grows the twiddle tables")
    grow_twiddles(nnh, 1 << 24)
    for rep in range(2):
        db.clear()
        b.run()
        nnh.synchronize()
        check_batch(db, ops, oracle, f"replay after growth {rep}")
    say("shutdown with the graph alive, then replay")
    nnh._check(nnh.lib().nnh_shutdown())
    db.clear()
    b.run()
    nnh.synchronize()
    check_batch(db, ops, oracle, "replay after shutdown")
    b.close()


def test_eager_batch_after_shutdown(nnh, oracle):
    """A batch without a graph, planned, then nnh_shutdown, then run: its
    launches read the tables the batch holds."""
    rng = np.random.default_rng(2025)
    sizes = [(int(2 ** rng.uniform(8, 13)), int(2 ** rng.uniform(8, 13))) for _ in range(24)]
    ops = make_ops(rng, sizes)
    db = DevBatch(nnh, ops)
    b = nnh.Batch(db.reqs, graph=False)
    nnh._check(nnh.lib().nnh_shutdown())
    say("eager batch after shutdown")
    b.run()
    nnh.synchronize()
    check_batch(db, ops, oracle, "eager run after shutdown")
    grow_twiddles(nnh, 1 << 20)
    db.clear()
    b.run()
    nnh.synchronize()
    check_batch(db, ops, oracle, "eager run after growth")
    b.close()


def test_async_products_on_two_streams(nnh, oracle):
    """Device-pointer products queued back to back on two caller streams
    with no synchronisation between them (each call's work waits for the
    previous call's in the shared workspace)."""
    rng = np.random.default_rng(2026)
    s1, s2 = nnh.Stream(), nnh.Stream()
    shapes = [(1 << 19, 1 << 19), (300000, 200000), (5000, 4000), (1 << 18, 1 << 17), (100, 7000), (70000, 70000)]
    ops = [(rand_limbs(rng, an), rand_limbs(rng, bn)) for an, bn in shapes]
    dev = []
    for a, b in ops:
        dev.append((nnh.DeviceWords(a.size, a), nnh.DeviceWords(b.size, b), nnh.DeviceWords(a.size + b.size)))
    nnh.synchronize()
    for rep in range(2):
        say(f"two-stream products, round {rep}")
        for i, ((a, b), (da, db, do)) in enumerate(zip(ops, dev)):
            st = s1 if (i + rep) % 2 == 0 else s2
            nnh.mul_device_async(do.ptr, da.ptr, a.size, db.ptr, b.size, st.ptr)
        s1.synchronize()
        s2.synchronize()
        for (a, b), (_, _, do) in zip(ops, dev):
            assert np.array_equal(do.download(), oracle.mul(a, b)), (a.size, b.size, rep)
            do.upload(np.zeros(a.size + b.size, dtype=np.uint64))
        nnh.synchronize()
    for t in dev:
        for x in t:
            x.free()
    s1.close()
    s2.close()


def test_injected_failure_enqueues_nothing(nnh, oracle):
    """nnh_debug_fail_next(1): the next call returns NNH_ERR_HIP and leaves
    its destination untouched; the call after it is exact."""
    rng = np.random.default_rng(2027)
    a, b = rand_limbs(rng, 5000), rand_limbs(rng, 3000)
    L = nnh.lib()
    assert L.nnh_debug_fail_next(1) == 0
    with pytest.raises(nnh.NNHError, match="injected failure"):
        nnh.mul(a, b)
    assert L.nnh_debug_fail_next(0) == 0
    assert np.array_equal(nnh.mul(a, b), oracle.mul(a, b))
    assert L.nnh_debug_fail_next(3) == 0
    assert L.nnh_debug_fail_next(0) == 3
