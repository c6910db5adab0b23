"""Write tests/golden/factorial.json: SHA-256 digests of n! as little-endian
64-bit limbs (normalised), computed by the system GMP (mpz_fac_ui through its
exported symbols), for the GPU factorial parity tests (reference
natural.h:795-839; test.cpp:35-39 computes 10^7!).

Run from the repository root:  python tests/golden/make_factorial_golden.py
"""
import ctypes
import ctypes.util
import hashlib
import json
import os
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
CASES = [10**6, 10**7]


class _Mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


def gmp_factorial_limbs(n):
    g = ctypes.CDLL(ctypes.util.find_library("gmp") or "libgmp.so.10")
    g.__gmpz_fac_ui.argtypes = [ctypes.c_void_p, ctypes.c_ulong]
    g.__gmpz_sizeinbase.restype = ctypes.c_size_t
    g.__gmpz_sizeinbase.argtypes = [ctypes.c_void_p, ctypes.c_int]
    g.__gmpz_export.restype = ctypes.c_void_p
    g.__gmpz_export.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int, ctypes.c_size_t,
                                ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
    z = _Mpz()
    g.__gmpz_init(ctypes.byref(z))
    g.__gmpz_fac_ui(ctypes.byref(z), n)
    words = (g.__gmpz_sizeinbase(ctypes.byref(z), 2) + 63) // 64
    out = np.zeros(words, dtype=np.uint64)
    cnt = ctypes.c_size_t(0)
    g.__gmpz_export(out.ctypes.data, ctypes.byref(cnt), -1, 8, 0, 0, ctypes.byref(z))
    g.__gmpz_clear(ctypes.byref(z))
    return out[:cnt.value]


def main():
    res = {}
    for n in CASES:
        t = time.time()
        w = gmp_factorial_limbs(n)
        res[str(n)] = {"words": int(w.size), "sha256": hashlib.sha256(w.astype("<u8").tobytes()).hexdigest(),
                       "top": hex(int(w[-1])), "low_nonzero_word": int(np.flatnonzero(w)[0])}
        print(n, w.size, f"{time.time() - t:.1f}s", flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "factorial.json"), "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
