"""Several host threads calling the library at once.  The reference's natural
is not thread-safe (its twiddle tables are global, ntt_info.h:141-145); the C
ABI serialises calls with an internal lock and keeps the error text per
thread, so concurrent callers get exact results and their own errors.
ctypes releases the GIL around each foreign call, so the calls do overlap."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
from conftest import rand_limbs, to_int

pytestmark = pytest.mark.gpu

# sizes below and above the 384-word threshold, one split-path shape
SHAPES = [(5, 3), (300, 200), (500, 400), (4096, 4096), (20000, 1000), (70000, 65000)]


def _worker(nnh, tid):
    rng = np.random.default_rng(1000 + tid)
    bad = []
    for it in range(3):
        for an, bn in SHAPES:
            a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
            A, B = to_int(a), to_int(b)
            if to_int(nnh.mul(a, b)) != A * B:
                bad.append(("mul", tid, it, an, bn))
            if to_int(nnh.sqr(b)) != B * B:
                bad.append(("sqr", tid, it, bn))
        a, b = rand_limbs(rng, 3000), rand_limbs(rng, 1100)
        q, r = nnh.divmod_(a, b)
        A, B = to_int(a), to_int(b)
        if (to_int(q), to_int(r)) != divmod(A, B):
            bad.append(("divmod", tid, it))
        if nnh.to_dec(b) != str(B):
            bad.append(("to_dec", tid, it))
    return bad


def test_concurrent_callers_exact(nnh):
    print("4 host threads x 3 rounds of mul/sqr/divmod/to_dec", flush=True)
    with ThreadPoolExecutor(4) as ex:
        results = list(ex.map(lambda t: _worker(nnh, t), range(4)))
    bad = [x for r in results for x in r]
    assert not bad, bad[:10]


def test_errors_stay_with_their_thread(nnh):
    """Failing calls (an output buffer too small for b^e, refused inside the
    locked section after the product ran) overlap with passing ones: each
    failing thread reads its own message (the word count differs per
    thread), the passing threads' products are exact and their error text
    stays empty."""
    import ctypes
    import threading

    L = nnh.lib()
    start = threading.Barrier(6)
    out = {}

    def failing(t):
        base = np.array([3], dtype=np.uint64)
        e = 5000 + 1000 * t
        buf = np.zeros(4, dtype=np.uint64)
        n = ctypes.c_uint64()
        start.wait()
        rc = L.nnh_pow(nnh._arr(buf), buf.size, ctypes.byref(n), nnh._arr(base), 1, e)
        need = -(-(3 ** e).bit_length() // 64)  # words of 3^e
        out[("fail", t)] = (rc, L.nnh_last_error().decode(), need)

    def passing(t):
        rng = np.random.default_rng(77 + t)
        a, b = rand_limbs(rng, 900), rand_limbs(rng, 800)
        start.wait()
        ok = to_int(nnh.mul(a, b)) == to_int(a) * to_int(b)
        out[("pass", t)] = (ok, L.nnh_last_error().decode())

    threads = [threading.Thread(target=failing, args=(t,)) for t in range(2)]
    threads += [threading.Thread(target=passing, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(60)
    assert len(out) == 6, sorted(out)
    for t in range(2):
        rc, msg, need = out[("fail", t)]
        assert rc != 0 and "too small" in msg and f"({need} words needed)" in msg, (t, rc, msg, need)
    for t in range(4):
        ok, msg = out[("pass", t)]
        assert ok and msg == "", (t, ok, msg)
