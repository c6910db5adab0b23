"""Transform handles (nnh_xf_*), the device counterparts of the reference's
mul_save_NTT_info / mul_load_NTT_info / sqr_save_NTT_info (natural.h:452-527),
against the oracle word for word: one kept transform reused by several
products, two kept transforms, squares of a kept transform, and products
longer than the transform, which NTT_info::save returns modulo B^L - 1
(ntt_info.h:237-248; the oracle's ntt_product runs the same rule).  Also the
reference's calc_e step (natural.h:852-856) with den1's transform kept, and
a handle surviving nnh_shutdown."""
import numpy as np
import pytest

from conftest import rand_limbs, to_int

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dev(nnh, x):
    return nnh.DeviceWords(x.size, x)


@pytest.mark.parametrize("scale,an,bns", [
    (8, 100, [150, 1, 156]),
    (10, 600, [400, 424, 1, 1024]),         # 1024: product longer than the transform (cyclic)
    (12, 4000, [96, 2000, 4096]),
    (14, 9000, [7000, 7385, 16000]),
    (20, 600000, [400000, 448577, 1 << 20]),
    (23, 1 << 22, [(1 << 22) - 5, 1 << 23]),
])
def test_kept_transform_products(nnh, oracle, scale, an, bns):
    rng = np.random.default_rng(scale)
    L = 1 << scale
    a = rand_limbs(rng, an)
    da = dev(nnh, a)
    xa = nnh.Transform(scale)
    xa.forward(da.ptr, an)
    for bn in bns:
        b = rand_limbs(rng, bn)
        db = dev(nnh, b)
        out = nnh.DeviceWords(max(an + bn, L + 2))
        n = xa.mul(out.ptr, db.ptr, bn)
        exp = oracle.ntt_product(a, b, scale)
        say(f"scale {scale}: {an} x {bn} -> {n} words ({'cyclic' if an + bn - 1 > L else 'exact'})")
        assert n == exp.size == (L if an + bn - 1 > L else an + bn)
        assert np.array_equal(out.download(n), exp), (scale, an, bn)
        db.free()
        out.free()
    # a square and a product of two kept transforms
    out = nnh.DeviceWords(max(2 * an, L + 2))
    n = xa.sqr(out.ptr)
    assert np.array_equal(out.download(n), oracle.ntt_product(a, a, scale))
    b = rand_limbs(rng, max(1, L - an))
    xb = nnh.Transform(scale)
    db = dev(nnh, b)
    xb.forward(db.ptr, b.size)
    n = xa.mul2(out.ptr, xb)
    assert np.array_equal(out.download(n), oracle.ntt_product(a, b, scale))
    for x in (xa, xb, da, db, out):
        x.close() if hasattr(x, "close") else x.free()


def test_calc_e_step_with_kept_transform(nnh):
    """natural.h:852-856: ni_den1 at scale get_NTT_scale(max(|den|, |num|),
    |den1|); num = den1 * num (mul_save: den1's transform kept), den =
    den1 * den (mul_load), both exact."""
    rng = np.random.default_rng(3)
    num, den, den1 = rand_limbs(rng, 30000), rand_limbs(rng, 31000), rand_limbs(rng, 29000)
    s = (max(den.size, num.size) + den1.size - 2).bit_length()
    x = nnh.Transform(s)
    d1, dn, dd = dev(nnh, den1), dev(nnh, num), dev(nnh, den)
    o1, o2 = nnh.DeviceWords(den1.size + num.size), nnh.DeviceWords(den1.size + den.size)
    x.forward(d1.ptr, den1.size)
    n1 = x.mul(o1.ptr, dn.ptr, num.size)
    n2 = x.mul(o2.ptr, dd.ptr, den.size)
    assert to_int(o1.download(n1)) == to_int(den1) * to_int(num)
    assert to_int(o2.download(n2)) == to_int(den1) * to_int(den)
    for y in (x, d1, dn, dd, o1, o2):
        y.close() if hasattr(y, "close") else y.free()


def test_handle_survives_shutdown(nnh, oracle):
    rng = np.random.default_rng(4)
    a, b = rand_limbs(rng, 3000), rand_limbs(rng, 5000)
    da, db = dev(nnh, a), dev(nnh, b)
    x = nnh.Transform(13)
    x.forward(da.ptr, a.size)
    nnh._check(nnh.lib().nnh_shutdown())
    out = nnh.DeviceWords(a.size + b.size)
    n = x.mul(out.ptr, db.ptr, b.size)
    assert np.array_equal(out.download(n), oracle.mul(a, b))
    with pytest.raises(nnh.NNHError):
        x.mul(out.ptr, db.ptr, 1 << 14)  # the operand must fit the transform
    for y in (x, da, db, out):
        y.close() if hasattr(y, "close") else y.free()
