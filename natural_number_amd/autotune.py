"""Autotuner of the transform's level split (the chunk size of the fused
chunk kernel and the split of the remaining levels into strided top passes),
per transform scale, for single products and for batch buckets separately.

The reference fixes its blocking (global levels down to 2^20-word chunks,
ntt_info.h:179-221); here the library times every candidate split on the
device and installs the fastest; the table is written next to the library
(plans_gfx950.txt), which loads it at nnh_init.  Single products are tuned
scale by scale (nnh_tune_plans: warm-up, interleaved rounds, a 1 % margin
over the formula).  Batch buckets run concurrently inside one batch, so
their splits are tuned on a whole batch workload (tune_batch_workload:
coordinate descent over the bucket scales, each candidate timed on the full
batch's graph).  Every split computes the same transform, so products are
bit-exact under any of them.
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
                       ("nnh_plan_candidates", [i32, ctypes.POINTER(PlanTime), i32, ctypes.POINTER(i32)]),
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


def candidates(scale):
    """[(chunk_bits, passes)] of the search space at `scale` (the formula's first)."""
    L = _sigs()
    cap = 64
    res = (PlanTime * cap)()
    n = ctypes.c_int()
    _check(L.nnh_plan_candidates(scale, res, cap, ctypes.byref(n)))
    return [(r.chunk_bits, list(r.pass_levels)[:r.n_passes]) for r in res[:n.value]]


def tune_batch_workload(reqs, scales=BATCH_SCALES, reps=3, rounds=2, margin=0.01, log=None):
    """Coordinate descent over the batch-bucket splits of `scales` on one
    batch workload (reqs: (dst, a, an, b, bn) device-pointer requests): for
    each scale in turn, every candidate split is timed on the whole batch
    (planned and captured as a graph, replayed `reps` times, the fastest of
    `rounds`) with the other scales' splits fixed, and kept when it beats the
    scale's current split by more than `margin`.  Returns
    {scale: {"split", "ms", "formula_ms"}} and the batch's ms before and after."""
    from . import Batch, Event, Stream, synchronize

    st = Stream()
    e0, e1 = Event(), Event()

    def time_batch():
        b = Batch(reqs, graph=True)
        best = None
        for _ in range(rounds):
            b.run(st.ptr)
            synchronize()
            e0.record(st)
            for _ in range(reps):
                b.run(st.ptr)
            e1.record(st)
            synchronize()
            t = e0.elapsed_ms(e1) / reps
            best = t if best is None else min(best, t)
        b.close()
        return best

    for _ in range(3):  # settle the clocks
        time_batch()
    start = time_batch()
    out = {}
    for s in scales:
        cands = candidates(s)
        if len(cands) < 2:
            continue
        cur_cb, cur_passes, tuned = plan(s, True)
        cur = (cur_cb, cur_passes)
        t_cur = time_batch()
        times = {}
        for cb, passes in cands:
            set_plan(s, True, cb, passes)
            times[(cb, tuple(passes))] = time_batch()
        (cb, passes), t_best = min(times.items(), key=lambda kv: kv[1])
        f_cb, f_passes = cands[0]
        if t_best < t_cur * (1 - margin) and (cb, list(passes)) != (f_cb, f_passes):
            set_plan(s, True, cb, list(passes))
            cur, t_cur = (cb, list(passes)), t_best
        elif t_best < t_cur * (1 - margin):
            set_plan(s, True, 0)  # the formula's split wins
            cur, t_cur = (f_cb, f_passes), t_best
        elif tuned:
            set_plan(s, True, cur_cb, cur_passes)
        else:
            set_plan(s, True, 0)
        out[s] = {"split": [cur[0]] + list(cur[1]), "ms": t_cur,
                  "formula_ms": times[(f_cb, tuple(f_passes))]}
        if log:
            log(f"batch workload s={s}: {out[s]['split']} {t_cur:.3f} ms (formula {out[s]['formula_ms']:.3f} ms)")
    end = time_batch()
    st.close()
    return {"scales": out, "batch_ms_before": start, "batch_ms_after": end}


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


def tune_all(single=SINGLE_SCALES, batch=(), log=None):
    """Tune scale by scale (isolated: single products, and batch buckets
    given in `batch`); returns {"single"|"batch": {scale: {"formula_ms",
    "best_ms", "best", "installed", "candidates"}}}."""
    out = {"single": {}, "batch": {}}
    for kind, scales in (("single", single), ("batch", batch)):
        for s in scales:
            r = tune(s, kind == "batch")
            best = min(r, key=lambda x: x["ms"])
            cb, passes, tuned = plan(s, kind == "batch")
            out[kind][s] = {"formula": [r[0]["chunk_bits"]] + r[0]["passes"], "formula_ms": r[0]["ms"],
                            "best": [best["chunk_bits"]] + best["passes"], "best_ms": best["ms"],
                            "installed": [cb] + passes if tuned else "formula", "candidates": len(r)}
            if log:
                log(f"tune {kind} s={s}: formula {out[kind][s]['formula']} {r[0]['ms']:.4f} ms, "
                    f"best {out[kind][s]['best']} {best['ms']:.4f} ms")
    return out
