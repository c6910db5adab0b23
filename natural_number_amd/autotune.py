"""Autotuner of the transform's level split (the chunk size of the fused
chunk kernel and the split of the remaining levels into strided top passes),
per transform scale, for single products and for batch buckets separately.

The reference fixes its blocking (global levels down to 2^20-word chunks,
ntt_info.h:179-221); here the library times every candidate split on the
device (nnh_tune_plans) and installs the fastest; the table is written next
to the library (plans_gfx950.txt), which loads it at nnh_init.  Every split
computes the same transform, so products are bit-exact under any of them.
"""
import ctypes

from . import _check, lib

SINGLE_SCALES = range(13, 27)   # products of 2^12 .. 2^25-word operands
BATCH_SCALES = range(13, 21)    # the buckets of batched products


class PlanTime(ctypes.Structure):
    _fields_ = [("chunk_bits", ctypes.c_int), ("n_passes", ctypes.c_int), ("pass_levels", ctypes.c_int * 4),
                ("ms", ctypes.c_float)]


def _sigs():
    L = lib()
    i32 = ctypes.c_int
    for name, args in [("nnh_tune_plans", [i32, i32, i32, ctypes.POINTER(PlanTime), i32, ctypes.POINTER(i32)]),
                       ("nnh_plan_set", [i32, i32, i32, i32, ctypes.POINTER(i32)]),
                       ("nnh_plan_get", [i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]),
                       ("nnh_plans_save", [ctypes.c_char_p]), ("nnh_plans_load", [ctypes.c_char_p])]:
        f = getattr(L, name)
        f.argtypes = args
        f.restype = i32
    return L


def default_reps(scale):
    return 20 if scale <= 20 else (8 if scale <= 24 else 4)


def tune(scale, batched=False, reps=None):
    """Time every candidate split at `scale`; the fastest is installed.
    Returns [{"chunk_bits", "passes", "ms"}], the formula's split first."""
    L = _sigs()
    cap = 64
    res = (PlanTime * cap)()
    n = ctypes.c_int()
    _check(L.nnh_tune_plans(scale, 1 if batched else 0, reps or default_reps(scale), res, cap, ctypes.byref(n)))
    return [{"chunk_bits": r.chunk_bits, "passes": list(r.pass_levels)[:r.n_passes], "ms": float(r.ms)}
            for r in res[:n.value]]


def plan(scale, batched=False):
    """(chunk_bits, passes, tuned) of the split the library uses."""
    L = _sigs()
    cb, np_, lev = ctypes.c_int(), ctypes.c_int(), (ctypes.c_int * 4)()
    rc = L.nnh_plan_get(scale, 1 if batched else 0, ctypes.byref(cb), ctypes.byref(np_), lev)
    if rc < 0 or rc > 1:
        _check(rc)
    return cb.value, list(lev)[:np_.value], bool(rc)


def set_plan(scale, batched, chunk_bits, passes=()):
    """Install a split (chunk_bits 0: back to the formula)."""
    L = _sigs()
    lev = (ctypes.c_int * 4)(*list(passes) + [0] * (4 - len(passes)))
    _check(L.nnh_plan_set(scale, 1 if batched else 0, chunk_bits, len(passes), lev))


def save(path):
    _check(_sigs().nnh_plans_save(str(path).encode()))


def load(path):
    _check(_sigs().nnh_plans_load(str(path).encode()))


def tune_all(single=SINGLE_SCALES, batch=BATCH_SCALES, log=None):
    """Tune every scale; returns {"single"|"batch": {scale: {"formula_ms",
    "best_ms", "best", "candidates"}}}."""
    out = {"single": {}, "batch": {}}
    for kind, scales in (("single", single), ("batch", batch)):
        for s in scales:
            r = tune(s, kind == "batch")
            best = min(r, key=lambda x: x["ms"])
            out[kind][s] = {"formula": [r[0]["chunk_bits"]] + r[0]["passes"], "formula_ms": r[0]["ms"],
                            "best": [best["chunk_bits"]] + best["passes"], "best_ms": best["ms"],
                            "candidates": len(r)}
            if log:
                log(f"tune {kind} s={s}: formula {out[kind][s]['formula']} {r[0]['ms']:.4f} ms, "
                    f"best {out[kind][s]['best']} {best['ms']:.4f} ms")
    return out
