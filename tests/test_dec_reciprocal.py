"""Decimal-base reciprocals derived one from the next (CPU, Python integers).

The reference's calc_dec_base (natural.h:909-923) takes, for every base
B_i = 10^(18 * 2^i), the reciprocal of D_i = B_i << n_shl_i to S_i = 2 n_i + 1
words (natural.h:575-617), which is floor(2^(64 S_i) / D_i).  The library's
decimal output runs that Newton ladder for the top base only and derives each
lower one as (B_i * rec_{i+1}) >> t_i, t_i = 64 (S_{i+1} - S_i) - n_shl_{i+1}
+ n_shl_i.  These tests check that formula against integer division, which
does not depend on either method.
"""
import pytest


def _words(x):
    return max(1, (x.bit_length() + 63) // 64)


def _n_shl(b):
    top = b >> (64 * (_words(b) - 1))
    return (64 - top.bit_length() + 1) & 63


def _bases(top):
    b = [10**18]
    for _ in range(top):
        b.append(b[-1] ** 2)
    return b


def _floor_rec(b):
    S = 2 * _words(b) + 1
    return (1 << (64 * S)) // (b << _n_shl(b)), S


def _shift(bi, bj):
    Si, Sj = 2 * _words(bi) + 1, 2 * _words(bj) + 1
    return 64 * (Sj - Si) - _n_shl(bj) + _n_shl(bi)


def test_derived_reciprocal_is_the_floor():
    B = _bases(13)
    for i in range(4, 12):
        exact, _ = _floor_rec(B[i])
        above, _ = _floor_rec(B[i + 1])
        assert (B[i] * above) >> _shift(B[i], B[i + 1]) == exact, i


def test_derived_from_an_underestimate_is_at_most_one_below():
    """An estimate one below its floor above gives the floor or one less."""
    B = _bases(12)
    for i in range(4, 11):
        exact, _ = _floor_rec(B[i])
        above, _ = _floor_rec(B[i + 1])
        got = (B[i] * (above - 1)) >> _shift(B[i], B[i + 1])
        assert exact - 1 <= got <= exact, i


def test_a_wrong_shift_fails():
    """A shift off by one bit either way is detected (the check is not vacuous)."""
    B = _bases(8)
    exact, _ = _floor_rec(B[5])
    above, _ = _floor_rec(B[6])
    t = _shift(B[5], B[6])
    for dt in (-1, 1):
        assert (B[5] * above) >> (t + dt) != exact


@pytest.mark.parametrize("i", [4, 7, 10])
def test_reciprocal_size_matches_the_reference_accuracy(i):
    """The floor has S_i + 1 - size(D_i) words, the size the ladder returns."""
    b = _bases(i)[i]
    r, S = _floor_rec(b)
    assert _words(r) == S + 1 - _words(b << _n_shl(b))
