"""Write tests/golden/e_digits.json: digit strings of e computed by the
reference recipe (calc_e over [0, 2^k), times 10^D, divided, printed in
decimal; reference test.cpp:12-19) for the GPU parity tests.

Pins, all independent of the GPU build:
  * the CPU oracle's digit string (oracle/liboracle.so nno_e_digits);
  * for D <= 10^5: mpmath's e to D digits (floor(e 10^D)), an independent
    evaluation of the constant;
  * for every D: the system GMP (exported mpz functions) dividing the
    oracle's calc_e numerator * 10^D by its denominator and printing the
    quotient, an independent division and radix conversion.
Only digests, lengths and the first/last digits are stored.

Run from the repository root:  python tests/golden/make_e_golden.py
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
CASES = [(14, 10**4), (17, 10**5), (18, 10**6), (21, 10**7), (24, 10**8)]


class _Mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


def _gmp():
    g = ctypes.CDLL(ctypes.util.find_library("gmp") or "libgmp.so.10")
    g.__gmpz_import.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                ctypes.c_size_t, ctypes.c_void_p]
    g.__gmpz_get_str.restype = ctypes.c_void_p
    g.__gmpz_get_str.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]
    g.__gmpz_ui_pow_ui.argtypes = [ctypes.c_void_p, ctypes.c_ulong, ctypes.c_ulong]
    return g


def gmp_e_digits(num, den, D):
    g = _gmp()
    zn, zd, zp, zq = _Mpz(), _Mpz(), _Mpz(), _Mpz()
    for z in (zn, zd, zp, zq):
        g.__gmpz_init(ctypes.byref(z))
    g.__gmpz_import(ctypes.byref(zn), num.size, -1, 8, 0, 0, num.ctypes.data)
    g.__gmpz_import(ctypes.byref(zd), den.size, -1, 8, 0, 0, den.ctypes.data)
    g.__gmpz_ui_pow_ui(ctypes.byref(zp), 10, D)
    g.__gmpz_mul(ctypes.byref(zn), ctypes.byref(zn), ctypes.byref(zp))
    g.__gmpz_tdiv_q(ctypes.byref(zq), ctypes.byref(zn), ctypes.byref(zd))
    buf = ctypes.create_string_buffer(D + 64)
    g.__gmpz_get_str(buf, 10, ctypes.byref(zq))
    s = buf.value
    for z in (zn, zd, zp, zq):
        g.__gmpz_clear(ctypes.byref(z))
    return s


def oracle():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    lib.nno_e_digits.restype = ctypes.c_uint64
    lib.nno_e_digits.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64]
    u64p = ctypes.POINTER(ctypes.c_uint64)
    lib.nno_calc_e.argtypes = [u64p, ctypes.c_uint64, u64p, ctypes.c_uint64, ctypes.c_int, u64p, u64p]
    return lib


def oracle_calc_e(lib, k):
    cap = (24 << k) // 64 + 64
    num = np.zeros(cap, dtype=np.uint64)
    den = np.zeros(cap, dtype=np.uint64)
    nn, dn = ctypes.c_uint64(), ctypes.c_uint64()
    u64p = ctypes.POINTER(ctypes.c_uint64)
    lib.nno_calc_e(num.ctypes.data_as(u64p), cap, den.ctypes.data_as(u64p), cap, k, ctypes.byref(nn),
                   ctypes.byref(dn))
    return num[:nn.value].copy(), den[:dn.value].copy()


def main():
    lib = oracle()
    out = {}
    for k, D in CASES:
        t = time.time()
        buf = ctypes.create_string_buffer(D + 64)
        n = lib.nno_e_digits(buf, D + 64, k, D)
        s = buf.value
        assert n == len(s) == D + 1, (n, len(s))
        t_oracle = time.time() - t
        num, den = oracle_calc_e(lib, k)
        g = gmp_e_digits(num, den, D)
        assert g == s, f"GMP and oracle disagree at k={k} D={D}"
        entry = {"k": k, "digits": D, "length": len(s), "sha256": hashlib.sha256(s).hexdigest(),
                 "head": s[:60].decode(), "tail": s[-60:].decode(), "oracle_seconds": round(t_oracle, 2)}
        if D <= 10**5:
            import mpmath
            mpmath.mp.dps = D + 30
            ref = mpmath.nstr(mpmath.e, D + 20, strip_zeros=False).replace(".", "")[:D + 1]
            assert ref.encode() == s, f"mpmath disagrees at D={D}"
            entry["mpmath_agrees"] = True
        out[f"e_k{k}_D{D}"] = entry
        print(k, D, entry["sha256"][:16], f"{t_oracle:.1f}s", flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "e_digits.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
