"""Write tests/golden/pi_digits.json: digests of the reference's pi recipe
(test.cpp:22-32: calc_pi over [1, 2^k + 1), sqrt_10005 to acc words,
division, decimal output) computed by the CPU oracle (oracle/liboracle.so,
nno_pi_digits), whose recipe the CPU tests pin against an exact-integer
restatement and mpmath.  Only digests, lengths and head/tail digits are
stored.

Run from the repository root:  python tests/golden/make_pi_golden.py
"""
import ctypes
import hashlib
import json
import os
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
CASES = [(14, 200000, 12000), (17, 1000000, 60000), (20, 10000000, 600000)]


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    lib.nno_pi_digits.restype = ctypes.c_uint64
    lib.nno_pi_digits.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
    out = {}
    for k, D, acc in CASES:
        buf = ctypes.create_string_buffer(D + 64)
        t = time.time()
        n = lib.nno_pi_digits(buf, D + 64, k, D, acc)
        dt = time.time() - t
        s = buf.raw[:n]
        assert s.startswith(b"31415926535897932384626433832795028841971693993751")
        out[f"pi_k{k}_D{D}_acc{acc}"] = {"k": k, "digits": D, "acc": acc, "length": n,
                                          "sha256": hashlib.sha256(s).hexdigest(), "head": s[:60].decode(),
                                          "tail": s[-60:].decode(), "oracle_seconds": round(dt, 2)}
        print(k, D, acc, n, f"{dt:.1f}s", flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "pi_digits.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
