"""Autotuner and learned cost model for the choice of multiplication arm.

Candidates are exact arms of libnnh (any choice gives the same bits):
schoolbook columns, one Karatsuba level, one transform for the whole
product, the reference's one-level split, and the chunked transform at two
piece scales.  The autotuner times every feasible arm on the GPU
(nnh_time_algo, a C-side loop of back-to-back launches); the cost model is
an FP32 MLP [log2 bits of the larger operand, log2 bits of the smaller,
log2 batch count] -> predicted log2 milliseconds per arm, trained and run by
the library's own HIP kernels (nnh_mlp_*).  Installed with
CostModel.install() (nnh_policy_set) or saved as a policy file
(save_policy) that the library loads at nnh_init, the model is the
library's dispatch policy: NNH_ALGO_AUTO, nnh_mul, the natural layer and
batches then take the arm it picks.  The reference's own policy (fixed
thresholds, natural.h:444-450) stays available as NNH_ALGO_REFERENCE.
"""
import ctypes
import math

import numpy as np

from . import (ALGO_KARATSUBA, ALGO_NTT, ALGO_NTT_CHUNK, ALGO_NTT_SPLIT, ALGO_REFERENCE, ALGO_SCHOOL, DeviceWords,
               _check, lib, policy_arm)

ARMS = ["school", "karatsuba", "ntt", "ntt_split", "ntt_chunk", "ntt_chunk+1"]
N_ARMS = len(ARMS)
N_IN = 3
MAX_WORDS = 1 << 24
TIME_CAP_MAC = 1 << 28     # schoolbook / Karatsuba products beyond this many word-products are not tried
MAX_PIECES = 4096          # chunked transform: at most this many pieces (one batch)


def _scale(a, b):
    return max(0, int(a + b - 2).bit_length())


def arm_call(arm, an, bn):
    """(algo, param) of arm index for operand sizes (words)."""
    big, small = max(an, bn), min(an, bn)
    smin = max(_scale(small, small), 8)
    return [(ALGO_SCHOOL, 0), (ALGO_KARATSUBA, 0), (ALGO_NTT, 0), (ALGO_NTT_SPLIT, 0),
            (ALGO_NTT_CHUNK, smin), (ALGO_NTT_CHUNK, smin + 1)][arm]


def feasible(an, bn):
    """Arms worth timing for this shape (all arms are exact; these caps only
    keep hopeless candidates out of the search)."""
    big, small = max(an, bn), min(an, bn)
    f = np.zeros(N_ARMS, dtype=np.float32)
    f[0] = 1.0 if big * small <= TIME_CAP_MAC else 0.0
    f[1] = 1.0 if small >= 2 and big * small <= TIME_CAP_MAC else 0.0
    f[2] = 1.0
    f[3] = 1.0
    for k, arm in ((4, 4), (5, 5)):
        sp = arm_call(arm, an, bn)[1]
        piece = (1 << sp) - small + 1
        pieces = -(-big // piece)
        f[k] = 1.0 if 2 <= pieces <= MAX_PIECES else 0.0
    return f


def library_feasible(an, bn):
    """The library's feasibility rule (nnh_arm_feasible; needs no device)."""
    f = (ctypes.c_float * N_ARMS)()
    _check(lib().nnh_arm_feasible(an, bn, f))
    return np.array(list(f), dtype=np.float32)


def library_arm_call(arm, an, bn):
    algo, param = ctypes.c_int(), ctypes.c_int()
    _check(lib().nnh_arm_call(arm, an, bn, ctypes.byref(algo), ctypes.byref(param)))
    return algo.value, param.value


def features(an, bn, count=1):
    big, small = max(an, bn), min(an, bn)
    return np.array([math.log2(64 * big) / 30.0, math.log2(64 * small) / 30.0, math.log2(count) / 8.0],
                    dtype=np.float32)


class Bench:
    """Random operands of the largest size, resident on the device; timed
    products point into them.  Batch timings place the batch's requests at
    different offsets (64 spare words, so even the largest operands of a
    batch have distinct B pointers) and their outputs side by side."""

    def __init__(self, seed=0x4E4E2E19, max_words=MAX_WORDS, out_words=None):
        rng = np.random.default_rng(seed)
        self.max = max_words
        self.cap = max_words + 64
        self.a = DeviceWords(self.cap, rng.integers(0, 2**64, size=self.cap, dtype=np.uint64))
        self.b = DeviceWords(self.cap, rng.integers(0, 2**64, size=self.cap, dtype=np.uint64))
        self.out_words = out_words or 8 * max_words
        self.out = DeviceWords(self.out_words)

    def time_batch(self, arm, an, bn, count, reps=None):
        """ms per product of a batch of `count` an x bn products, every request
        on `arm` (nnh_time_batch: planned once, replayed as a graph)."""
        work = count * (an + bn)
        if reps is None:
            reps = 8 if work < (1 << 20) else (3 if work < (1 << 24) else 1)
        ms = ctypes.c_float()
        _check(lib().nnh_time_batch(self.out.ptr, self.out_words, self.a.ptr, self.cap, an, self.b.ptr, self.cap, bn,
                                    arm, count, reps, ctypes.byref(ms)))
        return float(ms.value)

    def time(self, algo, param, an, bn, count=1, reps=None):
        if reps is None:
            reps = 8 if an * bn < (1 << 22) else (3 if an + bn < (1 << 22) else 1)
        ms = ctypes.c_float()
        _check(lib().nnh_time_algo(self.out.ptr, self.a.ptr, an, self.b.ptr, bn, algo, param, count, reps,
                                   ctypes.byref(ms)))
        return float(ms.value)

    def time_arm(self, arm, an, bn, count=1):
        algo, param = arm_call(arm, an, bn)
        return self.time(algo, param, an, bn, count)

    def close(self):
        for x in (self.a, self.b, self.out):
            x.free()


def grid(n_side=20, lo_bits=10, hi_bits=30, counts=(1, 8, 64)):
    """Training grid: log-uniform (larger, smaller) operand sizes and batch
    counts (1: a product on its own; 8, 64: that many products in one batch)."""
    pts = []
    for i in range(n_side):
        for j in range(i + 1):
            big = max(1, math.ceil(2 ** (lo_bits + (hi_bits - lo_bits) * i / (n_side - 1)) / 64))
            small = max(1, math.ceil(2 ** (lo_bits + (hi_bits - lo_bits) * j / (n_side - 1)) / 64))
            for c in counts:
                if c * big <= (1 << 26):
                    pts.append((big, small, c))
    return pts


def time_point(bench, arm, an, bn, count):
    """ms per product of arm at this grid point: on its own (count 1,
    nnh_time_algo) or inside a batch of `count` (nnh_time_batch)."""
    if count == 1:
        algo, param = arm_call(arm, an, bn)
        return bench.time(algo, param, an, bn)
    return bench.time_batch(arm, an, bn, count)


def collect(bench, points, log=None, repeats=2):
    """Measured log2(ms per product) for every feasible arm of every point:
    the fastest of `repeats` timings, so that one disturbed measurement
    does not bend the fitted landscape around its grid point."""
    X = np.zeros((len(points), N_IN), dtype=np.float32)
    T = np.zeros((len(points), N_ARMS), dtype=np.float32)
    M = np.zeros((len(points), N_ARMS), dtype=np.float32)
    for k, (an, bn, c) in enumerate(points):
        X[k] = features(an, bn, c)
        f = feasible(an, bn)
        for arm in range(N_ARMS):
            if f[arm]:
                ms = min(time_point(bench, arm, an, bn, c) for _ in range(repeats))
                T[k, arm] = math.log2(max(ms, 1e-6))
                M[k, arm] = 1.0
        if log and k % 50 == 0:
            log(f"collect {k}/{len(points)}")
    return X, T, M


class CostModel:
    """The FP32 MLP in device memory (nnh_mlp_*)."""

    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        _check(lib().nnh_mlp_create(N_IN, N_ARMS, seed, ctypes.byref(h)))
        self.h = h
        self.n_params = lib().nnh_mlp_num_params(h)

    def params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        _check(lib().nnh_mlp_get_params(self.h, p.ctypes.data, p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().nnh_mlp_set_params(self.h, p.ctypes.data, p.size))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.empty((X.shape[0], N_ARMS), dtype=np.float32)
        _check(lib().nnh_mlp_forward(self.h, X.ctypes.data, X.shape[0], y.ctypes.data))
        return y

    def gradients(self, X, T, M):
        X, T, M = (np.ascontiguousarray(v, dtype=np.float32) for v in (X, T, M))
        g = np.empty(self.n_params, dtype=np.float32)
        loss = ctypes.c_float()
        _check(lib().nnh_mlp_gradients(self.h, X.ctypes.data, T.ctypes.data, M.ctypes.data, X.shape[0],
                                       g.ctypes.data, ctypes.byref(loss)))
        return g, float(loss.value)

    def train(self, X, T, M, steps=3000, lr=3e-3):
        X, T, M = (np.ascontiguousarray(v, dtype=np.float32) for v in (X, T, M))
        loss = ctypes.c_float()
        _check(lib().nnh_mlp_train(self.h, X.ctypes.data, T.ctypes.data, M.ctypes.data, X.shape[0], steps,
                                   ctypes.c_float(lr), ctypes.byref(loss)))
        return float(loss.value)

    def choose(self, X, F):
        X, F = (np.ascontiguousarray(v, dtype=np.float32) for v in (X, F))
        ch = np.empty(X.shape[0], dtype=np.int32)
        _check(lib().nnh_mlp_choose(self.h, X.ctypes.data, F.ctypes.data, X.shape[0], ch.ctypes.data))
        return ch

    def install(self):
        """Make this network the library's dispatch policy (nnh_policy_set)."""
        _check(lib().nnh_policy_set(self.h))

    def close(self):
        if self.h:
            lib().nnh_mlp_destroy(self.h)
            self.h = None


POLICY_HEADER = "nnh-policy 1 {n_in} {n_out} {n}"


def save_policy(model, path):
    """Write the network as a policy file (nnh_policy_load's format)."""
    p = model.params()
    with open(path, "w") as f:
        f.write(POLICY_HEADER.format(n_in=N_IN, n_out=N_ARMS, n=p.size) + "\n")
        for k in range(0, p.size, 8):
            f.write(" ".join(f"{v:.9g}" for v in p[k:k + 8]) + "\n")


def load_policy_params(path):
    with open(path) as f:
        head = f.readline().split()
        assert head[:2] == ["nnh-policy", "1"], "not a policy file"
        n = int(head[4])
        p = np.array(f.read().split(), dtype=np.float32)
    assert p.size == n
    return p


def reference_forward(p, X, n_in=N_IN, n_out=N_ARMS, H=64):
    """Float64 NumPy restatement of the MLP (gradient checks only)."""
    o = 0
    W1 = p[o:o + H * n_in].reshape(H, n_in); o += H * n_in
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + n_out * H].reshape(n_out, H); o += n_out * H
    b3 = p[o:o + n_out]
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3.T + b3


def masked_mse(p, X, T, M):
    y = reference_forward(p, X)
    return float(((y - T) ** 2 * M).sum() / M.sum())


def evaluate(bench, model, requests, log=None):
    """Exhaustive autotune of every held-out request (the judge) against the
    arm the library's dispatch takes with the model installed as its policy
    (nnh_policy_arm), the model evaluated exactly per request on the device
    (nnh_mlp_choose), and the reference policy (NNH_ALGO_REFERENCE)."""
    model.install()
    X = np.stack([features(a, b, 1) for a, b in requests])
    F = np.stack([feasible(a, b) for a, b in requests])
    exact = model.choose(X, F)
    t_best = t_policy = t_exact = t_ref = 0.0
    regrets = []
    hits = 0
    for k, (an, bn) in enumerate(requests):
        times = {}
        for arm in range(N_ARMS):
            if F[k, arm]:
                times[arm] = bench.time_arm(arm, an, bn)
        best = min(times.values())
        arm, _, _ = policy_arm(an, bn, 1)
        # sizes outside the policy table take the reference's rule
        tp = times[arm] if arm >= 0 else bench.time(ALGO_REFERENCE, 0, an, bn)
        ta = bench.time(ALGO_REFERENCE, 0, an, bn)
        t_best += best
        t_policy += tp
        t_exact += times[int(exact[k])]
        t_ref += ta
        regrets.append(tp / best - 1.0)
        hits += tp == best
        if log and k % 500 == 0:
            log(f"evaluate {k}/{len(requests)}")
    r = np.array(regrets)
    return {
        "requests": len(requests),
        "total_ms_exhaustive_best": t_best,
        "total_ms_model": t_policy,
        "total_ms_model_exact_inference": t_exact,
        "total_ms_reference_policy": t_ref,
        "regret_total": t_policy / t_best - 1.0,
        "regret_total_exact_inference": t_exact / t_best - 1.0,
        "regret_mean": float(r.mean()),
        "regret_p90": float(np.quantile(r, 0.9)),
        "picked_best_fraction": hits / len(requests),
        "gain_over_reference_policy": t_ref / t_policy,
        "reference_policy_regret_total": t_ref / t_best - 1.0,
    }
