"""Write tests/golden/products.json: SHA-256 digests (and a few sampled
words) of the products the GPU parity tests check at full size.

Inputs are splitmix64 streams (natural_number_amd.splitmix64) so only the
seeds and sizes are stored.  Each product is computed by two independent
exact multipliers when available: the system GMP (called through its
exported mpz functions) and the CPU oracle (oracle/liboracle.so, a
restatement of the reference's three-prime NTT path).  They must agree.

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes
import ctypes.util
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from natural_number_amd import splitmix64  # noqa: E402

CASES = [
    # name, a (seed, words), b (seed, words)
    ("c1_2^20x2^20", (1, 1 << 14), (2, 1 << 14)),
    ("c2_2^30x2^30", (3, 1 << 24), (4, 1 << 24)),
    ("c2_2^30x2^22", (5, 1 << 24), (6, 1 << 16)),
    ("c2_2^28x2^28", (7, 1 << 22), (8, 1 << 22)),
    ("mid_2^25x2^23", (9, 1 << 19), (10, 1 << 17)),
]

_u64p = ctypes.POINTER(ctypes.c_uint64)


class _Mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


def gmp_mul(a, b):
    path = ctypes.util.find_library("gmp") or "libgmp.so.10"
    g = ctypes.CDLL(path)
    za, zb, zc = _Mpz(), _Mpz(), _Mpz()
    for z in (za, zb, zc):
        g.__gmpz_init(ctypes.byref(z))
    imp = g.__gmpz_import
    imp.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t,
                    ctypes.c_void_p]
    imp(ctypes.byref(za), a.size, -1, 8, 0, 0, a.ctypes.data)
    imp(ctypes.byref(zb), b.size, -1, 8, 0, 0, b.ctypes.data)
    g.__gmpz_mul(ctypes.byref(zc), ctypes.byref(za), ctypes.byref(zb))
    out = np.zeros(a.size + b.size, dtype=np.uint64)
    cnt = ctypes.c_size_t(0)
    exp = g.__gmpz_export
    exp.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                    ctypes.c_size_t, ctypes.c_void_p]
    exp.restype = ctypes.c_void_p
    exp(out.ctypes.data, ctypes.byref(cnt), -1, 8, 0, 0, ctypes.byref(zc))
    for z in (za, zb, zc):
        g.__gmpz_clear(ctypes.byref(z))
    return out


def oracle_mul(a, b):
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    lib.nno_mul.restype = ctypes.c_uint64
    out = np.zeros(a.size + b.size, dtype=np.uint64)
    lib.nno_mul(out.ctypes.data_as(_u64p), a.ctypes.data_as(_u64p), ctypes.c_uint64(a.size),
                b.ctypes.data_as(_u64p), ctypes.c_uint64(b.size))
    return out


def digest(w):
    return hashlib.sha256(np.ascontiguousarray(w, dtype="<u8").tobytes()).hexdigest()


def main():
    res = {}
    for name, (sa, na), (sb, nb) in CASES:
        a, b = splitmix64(sa, na), splitmix64(sb, nb)
        t = time.time()
        g = gmp_mul(a, b)
        tg = time.time() - t
        t = time.time()
        o = oracle_mul(a, b)
        to = time.time() - t
        if not np.array_equal(g, o):
            raise SystemExit(f"{name}: GMP and oracle disagree")
        idx = [0, 1, (na + nb) // 2, na + nb - 2, na + nb - 1]
        res[name] = {
            "a": {"seed": sa, "words": na},
            "b": {"seed": sb, "words": nb},
            "sha256": digest(g),
            "sample": {str(i): hex(int(g[i])) for i in idx},
        }
        print(f"{name}: gmp {tg:.2f}s oracle {to:.2f}s {res[name]['sha256'][:16]}", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "products.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
