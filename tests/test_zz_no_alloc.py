"""The allocation-free product entry point (nnh_mul_device_noalloc) against
the allocating one (nnh_mul_device_async) and the oracle: word for word on
consecutive calls whose sizes shrink (schoolbook, direct transform and the
one-level split, which use the shared scratch array differently), and the
size limit."""
import numpy as np
import pytest

from conftest import rand_limbs

pytestmark = pytest.mark.gpu

LIMIT = 1 << 20  # words of the largest allocation-free product


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev_ops(nnh):
    rng = np.random.default_rng(0x4E4E)
    # largest first, then smaller inputs: each call must not depend on what
    # the previous one left in the shared array
    shapes = [(600000, 424288), (700000, 1000), (40000, 30000), (5000, 300), (384, 384), (3000, 3000), (17, 5)]
    ops = []
    for an, bn in shapes:
        assert an + bn <= LIMIT
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        ops.append((a, b, nnh.DeviceWords(an, a), nnh.DeviceWords(bn, b), nnh.DeviceWords(an + bn),
                    nnh.DeviceWords(an + bn)))
    nnh.synchronize()
    yield ops
    for t in ops:
        for x in t[2:]:
            x.free()


def test_noalloc_matches_allocating_entry_point(nnh, oracle, dev_ops):
    st = nnh.Stream()
    for rnd in range(2):
        for a, b, da, db, d_noalloc, d_alloc in dev_ops:
            say(f"round {rnd}: {a.size} x {b.size} words")
            d_noalloc.upload(np.zeros(a.size + b.size, dtype=np.uint64))
            d_alloc.upload(np.zeros(a.size + b.size, dtype=np.uint64))
            nnh.mul_device_noalloc(d_noalloc.ptr, da.ptr, a.size, db.ptr, b.size, st.ptr)
            nnh.mul_device_async(d_alloc.ptr, da.ptr, a.size, db.ptr, b.size, st.ptr)
            st.synchronize()
            got = d_noalloc.download()
            assert np.array_equal(got, d_alloc.download()), (rnd, a.size, b.size)
            if rnd == 0:
                assert np.array_equal(got, oracle.mul(a, b)), (a.size, b.size)
    st.close()


def test_noalloc_two_calls_in_a_row_second_smaller(nnh, dev_ops):
    """Two allocation-free calls queued back to back, the second smaller,
    then both outputs read: each equals the allocating entry point's."""
    (a1, b1, da1, db1, o1, r1), (a2, b2, da2, db2, o2, r2) = dev_ops[0], dev_ops[2]
    for o in (o1, o2):
        o.upload(np.zeros(o.n, dtype=np.uint64))
    nnh.mul_device_noalloc(o1.ptr, da1.ptr, a1.size, db1.ptr, b1.size)
    nnh.mul_device_noalloc(o2.ptr, da2.ptr, a2.size, db2.ptr, b2.size)
    nnh.mul_device_async(r1.ptr, da1.ptr, a1.size, db1.ptr, b1.size)
    nnh.mul_device_async(r2.ptr, da2.ptr, a2.size, db2.ptr, b2.size)
    nnh.synchronize()
    assert np.array_equal(o1.download(), r1.download())
    assert np.array_equal(o2.download(), r2.download())


def test_noalloc_rejects_products_beyond_its_limit(nnh, dev_ops):
    a, b, da, db, o, _ = dev_ops[0]
    # buffers large enough for the refused product, so nothing would be out of bounds even if it ran
    big_in = nnh.DeviceWords(LIMIT)
    big_out = nnh.DeviceWords(LIMIT + 1)
    with pytest.raises(nnh.NNHError, match="allocation-free limit"):
        nnh.mul_device_noalloc(big_out.ptr, big_in.ptr, LIMIT - 10, db.ptr, 11)
    big_in.free()
    big_out.free()
    # the library keeps working after the refusal
    o.upload(np.zeros(a.size + b.size, dtype=np.uint64))
    nnh.mul_device_noalloc(o.ptr, da.ptr, a.size, db.ptr, b.size)
    nnh.synchronize()
    exp = nnh.DeviceWords(a.size + b.size)
    nnh.mul_device_async(exp.ptr, da.ptr, a.size, db.ptr, b.size)
    nnh.synchronize()
    assert np.array_equal(o.download(), exp.download())
    exp.free()
