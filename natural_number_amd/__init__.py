"""natural_number_amd — MI355X three-prime NTT multiply of large naturals.

The compute lives in libnnh.so (HIP kernels for gfx950 behind the C ABI in
include/nnh.h).  This module is a thin ctypes binding used by the tests and
the benchmark; it never computes a product itself and raises NNHError when
the library is missing or a call fails.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "NNHError", "lib", "mul", "sqr", "mul_cyclic", "transform_forward", "alpha",
    "mul_transformed", "mul_device_async", "sqr_device_async", "threshold_words", "kernel_timing", "kernel_times",
    "to_limbs", "from_limbs", "splitmix64", "Req", "Batch", "mul_batch",
    "DeviceWords", "Stream", "Event", "synchronize",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnnh.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)


class NNHError(RuntimeError):
    pass


class Req(ctypes.Structure):
    """struct nnh_req: device pointers and word counts of one product."""
    _fields_ = [("dst", ctypes.c_void_p), ("a", ctypes.c_void_p), ("an", ctypes.c_uint64),
                ("b", ctypes.c_void_p), ("bn", ctypes.c_uint64)]


def lib():
    """Load libnnh.so (built in-tree by natural_number_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NNHError(f"{LIB_PATH} is not built; run `python -m natural_number_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    u64, vp, i32 = ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int
    sigs = {
        "nnh_init": ([i32], i32),
        "nnh_shutdown": ([], i32),
        "nnh_debug_fail_next": ([i32], i32),
        "nnh_mul": ([_u64p, _u64p, u64, _u64p, u64], i32),
        "nnh_sqr": ([_u64p, _u64p, u64], i32),
        "nnh_mul_device_async": ([vp, vp, u64, vp, u64, vp], i32),
        "nnh_sqr_device_async": ([vp, vp, u64, vp], i32),
        "nnh_mul_device_noalloc": ([vp, vp, u64, vp, u64, vp], i32),
        "nnh_mul_cyclic": ([_u64p, _u64p, u64, _u64p, u64, i32], i32),
        "nnh_transform_forward": ([_u64p, _u64p, u64, i32], i32),
        "nnh_alpha": ([i32, _u64p], i32),
        "nnh_mul_transformed": ([_u64p, _u64p, u64, _u64p, u64, i32], i32),
        "nnh_threshold_words": ([], u64),
        "nnh_kernel_timing": ([i32], i32),
        "nnh_kernel_times": ([i32, ctypes.POINTER(ctypes.c_double), _u64p, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_char_p)], i32),
        "nnh_last_error": ([], ctypes.c_char_p),
        "nnh_batch_create": ([ctypes.POINTER(Req), u64, i32, ctypes.POINTER(vp)], i32),
        "nnh_batch_run": ([vp, vp], i32),
        "nnh_batch_destroy": ([vp], i32),
        "nnh_mul_batch": ([ctypes.POINTER(Req), u64, vp], i32),
        "nnh_mul_algo": ([_u64p, _u64p, u64, _u64p, u64, i32, i32], i32),
        "nnh_mul_algo_device_async": ([vp, vp, u64, vp, u64, i32, i32, vp], i32),
        "nnh_time_algo": ([vp, vp, u64, vp, u64, i32, i32, i32, i32, ctypes.POINTER(ctypes.c_float)], i32),
        "nnh_time_batch": ([vp, u64, vp, u64, u64, vp, u64, u64, i32, i32, i32, ctypes.POINTER(ctypes.c_float)], i32),
        "nnh_butterfly_peak": ([i32, u64, ctypes.POINTER(ctypes.c_double)], i32),
        "nnh_transform_plan": ([i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)], i32),
        "nnh_mlp_create": ([i32, i32, u64, ctypes.POINTER(vp)], i32),
        "nnh_mlp_destroy": ([vp], i32),
        "nnh_mlp_num_params": ([vp], i32),
        "nnh_mlp_set_params": ([vp, vp, u64], i32),
        "nnh_mlp_get_params": ([vp, vp, u64], i32),
        "nnh_mlp_forward": ([vp, vp, u64, vp], i32),
        "nnh_mlp_gradients": ([vp, vp, vp, vp, u64, vp, ctypes.POINTER(ctypes.c_float)], i32),
        "nnh_mlp_train": ([vp, vp, vp, vp, u64, i32, ctypes.c_float, ctypes.POINTER(ctypes.c_float)], i32),
        "nnh_mlp_choose": ([vp, vp, vp, u64, vp], i32),
        "nnh_xf_create": ([i32, ctypes.POINTER(vp)], i32),
        "nnh_xf_destroy": ([vp], i32),
        "nnh_xf_forward": ([vp, vp, u64, vp], i32),
        "nnh_xf_mul": ([vp, _u64p, vp, vp, u64, vp], i32),
        "nnh_xf_mul2": ([vp, _u64p, vp, vp, vp], i32),
        "nnh_xf_sqr": ([vp, _u64p, vp, vp], i32),
        "nnh_policy_set": ([vp], i32),
        "nnh_policy_load": ([ctypes.c_char_p], i32),
        "nnh_policy_active": ([], i32),
        "nnh_policy_arm": ([u64, u64, u64, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)], i32),
        "nnh_arm_feasible": ([u64, u64, ctypes.POINTER(ctypes.c_float)], i32),
        "nnh_arm_call": ([i32, u64, u64, ctypes.POINTER(i32), ctypes.POINTER(i32)], i32),
        "nnh_device_alloc": ([ctypes.POINTER(vp), u64], i32),
        "nnh_device_free": ([vp], i32),
        "nnh_copy_to_device": ([vp, vp, u64], i32),
        "nnh_copy_to_host": ([vp, vp, u64], i32),
        "nnh_copy_device": ([vp, vp, u64, vp], i32),
        "nnh_stream_create": ([ctypes.POINTER(vp)], i32),
        "nnh_stream_destroy": ([vp], i32),
        "nnh_synchronize": ([vp], i32),
        "nnh_event_create": ([ctypes.POINTER(vp)], i32),
        "nnh_event_record": ([vp, vp], i32),
        "nnh_event_elapsed_ms": ([vp, vp, ctypes.POINTER(ctypes.c_float)], i32),
        "nnh_event_destroy": ([vp], i32),
        "nnh_divmod": ([_u64p, u64, _u64p, _u64p, u64, _u64p, _u64p, u64, _u64p, u64], i32),
        "nnh_pow": ([_u64p, u64, _u64p, _u64p, u64, u64], i32),
        "nnh_to_dec": ([ctypes.c_char_p, u64, _u64p, _u64p, u64], i32),
        "nnh_calc_e": ([_u64p, u64, _u64p, _u64p, u64, _u64p, i32], i32),
        "nnh_factorial": ([_u64p, u64, _u64p, u64], i32),
        "nnh_calc_pi": ([_u64p, _u64p, _u64p, u64, i32, _u64p], i32),
        "nnh_pi_digits": ([ctypes.c_char_p, u64, _u64p, i32, u64, u64, ctypes.POINTER(ctypes.c_double)], i32),
        "nnh_e_digits": ([ctypes.c_char_p, u64, _u64p, i32, u64, ctypes.POINTER(ctypes.c_double)], i32),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise NNHError(f"nnh status {rc}: {lib().nnh_last_error().decode(errors='replace')}")


def _arr(x):
    return x.ctypes.data_as(_u64p)


def _limbs(x):
    a = np.ascontiguousarray(x, dtype=np.uint64)
    if a.ndim != 1 or a.size == 0:
        raise ValueError("operands are non-empty 1-D uint64 limb arrays")
    return a


def mul(a, b):
    """a * b as an+bn little-endian words (not normalised)."""
    a, b = _limbs(a), _limbs(b)
    out = np.empty(a.size + b.size, dtype=np.uint64)
    _check(lib().nnh_mul(_arr(out), _arr(a), a.size, _arr(b), b.size))
    return out


ALGO_AUTO, ALGO_SCHOOL, ALGO_KARATSUBA, ALGO_NTT, ALGO_NTT_SPLIT, ALGO_NTT_CHUNK, ALGO_REFERENCE, ALGO_MODEL = range(8)
ALGO_NAMES = ["auto", "school", "karatsuba", "ntt", "ntt_split", "ntt_chunk", "reference", "model"]
BATCH_GRAPH, BATCH_REFERENCE = 1, 8


def mul_algo(a, b, algo, param=0):
    """a * b through one explicit algorithm arm (exact whichever)."""
    a, b = _limbs(a), _limbs(b)
    out = np.empty(a.size + b.size, dtype=np.uint64)
    _check(lib().nnh_mul_algo(_arr(out), _arr(a), a.size, _arr(b), b.size, algo, param))
    return out


def mul_algo_device_async(d_out, d_a, an, d_b, bn, algo, param=0, stream=None):
    _check(lib().nnh_mul_algo_device_async(d_out, d_a, an, d_b, bn, algo, param, stream))


def sqr(a):
    a = _limbs(a)
    out = np.empty(2 * a.size, dtype=np.uint64)
    _check(lib().nnh_sqr(_arr(out), _arr(a), a.size))
    return out


def mul_cyclic(a, b, scale):
    """Product through one transform of 2^scale words; mod 2^(64*2^scale)-1
    (2^scale words) when longer than the transform, else exact."""
    a, b = _limbs(a), _limbs(b)
    L = 1 << scale
    n = L if a.size + b.size - 1 > L else a.size + b.size
    out = np.empty(max(n, a.size + b.size), dtype=np.uint64)
    _check(lib().nnh_mul_cyclic(_arr(out), _arr(a), a.size, _arr(b), b.size, scale))
    return out[:n]


def transform_forward(a, scale):
    a = _limbs(a)
    out = np.empty(3 << scale, dtype=np.uint64)
    _check(lib().nnh_transform_forward(_arr(out), _arr(a), a.size, scale))
    return out


def alpha(scale):
    out = np.empty(3, dtype=np.uint64)
    _check(lib().nnh_alpha(scale, _arr(out)))
    return [int(x) for x in out]


def mul_transformed(fa, an, b, scale):
    fa, b = _limbs(fa), _limbs(b)
    out = np.empty(an + b.size, dtype=np.uint64)
    _check(lib().nnh_mul_transformed(_arr(out), _arr(fa), an, _arr(b), b.size, scale))
    return out


def mul_device_async(d_out, d_a, an, d_b, bn, stream=None):
    """Device-pointer product (pointers as ints), queued on `stream`."""
    _check(lib().nnh_mul_device_async(d_out, d_a, an, d_b, bn, stream))


def mul_device_noalloc(d_out, d_a, an, d_b, bn, stream=None):
    """Allocation-free product (an + bn <= 2^20 words), device pointers."""
    _check(lib().nnh_mul_device_noalloc(d_out, d_a, an, d_b, bn, stream))


def sqr_device_async(d_out, d_a, an, stream=None):
    _check(lib().nnh_sqr_device_async(d_out, d_a, an, stream))


def kernel_timing(enable):
    _check(lib().nnh_kernel_timing(1 if enable else 0))


def transform_plan(scale):
    """(chunk_bits, [levels of each top pass, forward order]) of a 2^scale-word transform."""
    cb, n = ctypes.c_int(), ctypes.c_int()
    lv = (ctypes.c_int * 4)()
    _check(lib().nnh_transform_plan(scale, ctypes.byref(cb), ctypes.byref(n), lv))
    return cb.value, [lv[i] for i in range(n.value)]


def butterfly_peak(prime, iters=256):
    """Register-resident butterflies per second of `prime`'s transform butterfly."""
    r = ctypes.c_double()
    _check(lib().nnh_butterfly_peak(prime, iters, ctypes.byref(r)))
    return r.value


def kernel_times():
    """{name: (total_ms, launches, algorithmic_bytes)} since timing was enabled."""
    res = {}
    for k in range(7):
        ms, n, by, nm = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_double(), ctypes.c_char_p()
        _check(lib().nnh_kernel_times(k, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(by), ctypes.byref(nm)))
        res[nm.value.decode()] = (ms.value, int(n.value), by.value)
    return res


class Batch:
    """A planned batch of device-pointer products (nnh_batch_create)."""

    def __init__(self, reqs, graph=True, reference=False):
        """graph: capture into a hipGraph; reference: plan with the
        reference's rule even when a cost-model policy is installed."""
        arr = (Req * len(reqs))(*[Req(*r) for r in reqs])
        h = ctypes.c_void_p()
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_REFERENCE if reference else 0)
        _check(lib().nnh_batch_create(arr, len(reqs), flags, ctypes.byref(h)))
        self.h = h

    def run(self, stream=None):
        _check(lib().nnh_batch_run(self.h, stream))

    def close(self):
        if self.h:
            _check(lib().nnh_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Transform:
    """A forward transform kept in device memory (nnh_xf_*): the library's
    counterpart of a kept NTT_info (mul_save / mul_load / sqr_save,
    natural.h:452-527)."""

    def __init__(self, scale):
        h = ctypes.c_void_p()
        _check(lib().nnh_xf_create(scale, ctypes.byref(h)))
        self.h = h
        self.scale = scale

    def forward(self, d_a, an, stream=None):
        _check(lib().nnh_xf_forward(self.h, d_a, an, stream))

    def mul(self, d_dst, d_b, bn, stream=None):
        """d_dst = a * b (save rule); returns the result's word count."""
        n = ctypes.c_uint64()
        _check(lib().nnh_xf_mul(d_dst, ctypes.byref(n), self.h, d_b, bn, stream))
        return n.value

    def mul2(self, d_dst, other, stream=None):
        n = ctypes.c_uint64()
        _check(lib().nnh_xf_mul2(d_dst, ctypes.byref(n), self.h, other.h, stream))
        return n.value

    def sqr(self, d_dst, stream=None):
        n = ctypes.c_uint64()
        _check(lib().nnh_xf_sqr(d_dst, ctypes.byref(n), self.h, stream))
        return n.value

    def close(self):
        if self.h:
            _check(lib().nnh_xf_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def policy_active():
    """True when a cost-model policy drives the library's dispatch."""
    return bool(lib().nnh_policy_active())


def policy_load(path):
    _check(lib().nnh_policy_load(str(path).encode()))


def policy_clear():
    _check(lib().nnh_policy_set(None))


def policy_arm(an, bn, count=1):
    """(arm, algo, param) the dispatch takes; arm -1 = the reference's rule."""
    arm, algo, param = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib().nnh_policy_arm(an, bn, count, ctypes.byref(arm), ctypes.byref(algo), ctypes.byref(param)))
    return arm.value, algo.value, param.value


def mul_batch(reqs, stream=None):
    """One-shot batch of (dst, a, an, b, bn) device-pointer products."""
    arr = (Req * len(reqs))(*[Req(*r) for r in reqs])
    _check(lib().nnh_mul_batch(arr, len(reqs), stream))


class DeviceWords:
    """A device buffer of 64-bit words owned through the library's runtime."""

    def __init__(self, n, host=None):
        self.n = int(n)
        p = ctypes.c_void_p()
        _check(lib().nnh_device_alloc(ctypes.byref(p), 8 * max(self.n, 1)))
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    def upload(self, host, offset=0):
        h = np.ascontiguousarray(host, dtype=np.uint64)
        _check(lib().nnh_copy_to_device(self.ptr + 8 * offset, h.ctypes.data, 8 * h.size))

    def download(self, n=None, offset=0):
        n = self.n - offset if n is None else n
        out = np.empty(n, dtype=np.uint64)
        _check(lib().nnh_copy_to_host(out.ctypes.data, self.ptr + 8 * offset, 8 * n))
        return out

    def at(self, offset):
        return self.ptr + 8 * offset

    def free(self):
        if self.ptr:
            _check(lib().nnh_device_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib().nnh_stream_create(ctypes.byref(p)))
        self.ptr = p.value

    def synchronize(self):
        _check(lib().nnh_synchronize(self.ptr))

    def close(self):
        if self.ptr:
            _check(lib().nnh_stream_destroy(self.ptr))
            self.ptr = None


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib().nnh_event_create(ctypes.byref(p)))
        self.ptr = p.value

    def record(self, stream=None):
        _check(lib().nnh_event_record(self.ptr, stream.ptr if isinstance(stream, Stream) else stream))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _check(lib().nnh_event_elapsed_ms(self.ptr, end.ptr, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.ptr:
            _check(lib().nnh_event_destroy(self.ptr))
            self.ptr = None


def synchronize(stream=None):
    _check(lib().nnh_synchronize(stream.ptr if isinstance(stream, Stream) else stream))


# ---- device-resident natural layer -------------------------------------------
def divmod_(a, b):
    """(a // b, a % b) by the reference's Newton division on the GPU
    (natural.h:749-769); both normalised uint64 limb arrays."""
    a, b = _limbs(a), _limbs(b)
    q = np.zeros(a.size + 1, dtype=np.uint64)
    r = np.zeros(b.size + 1, dtype=np.uint64)
    qn, rn = ctypes.c_uint64(), ctypes.c_uint64()
    _check(lib().nnh_divmod(_arr(q), q.size, ctypes.byref(qn), _arr(r), r.size, ctypes.byref(rn), _arr(a), a.size,
                            _arr(b), b.size))
    return q[:qn.value].copy(), r[:rn.value].copy()


def pow_(b, e):
    """b^e (natural.h:554-564), normalised limbs."""
    b = _limbs(b)
    bits = max(1, int(np.uint64(b[-1])).bit_length() + 64 * (b.size - 1))
    cap = (bits * max(e, 1)) // 64 + 2
    out = np.zeros(cap, dtype=np.uint64)
    n = ctypes.c_uint64()
    _check(lib().nnh_pow(_arr(out), cap, ctypes.byref(n), _arr(b), b.size, e))
    return out[:n.value].copy()


def to_dec(a):
    """Decimal string of a (natural.h:1020-1039)."""
    a = _limbs(a)
    cap = int(a.size * 64 * 0.30103) + 64
    buf = ctypes.create_string_buffer(cap)
    n = ctypes.c_uint64()
    _check(lib().nnh_to_dec(buf, cap, ctypes.byref(n), _arr(a), a.size))
    return buf.raw[:n.value].decode()


def calc_e(k):
    """(num, den) of calc_e(0, 2^k) (natural.h:841-857)."""
    cap = (24 << k) // 64 + 64 if k <= 24 else (k << k) // 64 + 64
    num = np.zeros(cap, dtype=np.uint64)
    den = np.zeros(cap, dtype=np.uint64)
    nn, dn = ctypes.c_uint64(), ctypes.c_uint64()
    _check(lib().nnh_calc_e(_arr(num), cap, ctypes.byref(nn), _arr(den), cap, ctypes.byref(dn), k))
    return num[:nn.value].copy(), den[:dn.value].copy()


def factorial(n):
    """n! (natural.h:795-839), normalised limbs."""
    import math
    bits = int(n * max(1.0, math.log2(max(n, 2)))) + 64
    cap = bits // 64 + 2
    out = np.zeros(cap, dtype=np.uint64)
    m = ctypes.c_uint64()
    _check(lib().nnh_factorial(_arr(out), cap, ctypes.byref(m), n))
    return out[:m.value].copy()


def calc_pi(k):
    """(num, den, numc) of calc_pi(1, 2^k + 1) (natural.h:874-902)."""
    cap = (340 << k) // 64 + 64
    bufs = [np.zeros(cap, dtype=np.uint64) for _ in range(3)]
    sizes = np.zeros(3, dtype=np.uint64)
    _check(lib().nnh_calc_pi(_arr(bufs[0]), _arr(bufs[1]), _arr(bufs[2]), cap, k, _arr(sizes)))
    return tuple(b[:int(n)].copy() for b, n in zip(bufs, sizes))


PI_STAGES = ["calc_pi", "sqrt_10005", "products", "div", "to_dec"]


def pi_digits_into(buf, k, digits, acc):
    """The reference's pi recipe (test.cpp:22-32) into a ctypes char buffer:
    (digit count, {stage: ms})."""
    n = ctypes.c_uint64()
    ms = (ctypes.c_double * len(PI_STAGES))()
    _check(lib().nnh_pi_digits(buf, len(buf), ctypes.byref(n), k, digits, acc, ms))
    return n.value, dict(zip(PI_STAGES, list(ms)))


def pi_digits(k, digits, acc):
    """The reference's pi recipe (test.cpp:22-32): (digit string, {stage: ms})."""
    buf = ctypes.create_string_buffer(digits + 64)
    n, st = pi_digits_into(buf, k, digits, acc)
    return buf.raw[:n].decode(), st


E_STAGES = ["calc_e", "pow", "mul", "div", "to_dec"]


def e_digits_into(buf, k, digits):
    """The reference's e recipe (test.cpp:12-19) into a ctypes char buffer:
    (digit count, {stage: ms})."""
    n = ctypes.c_uint64()
    ms = (ctypes.c_double * len(E_STAGES))()
    _check(lib().nnh_e_digits(buf, len(buf), ctypes.byref(n), k, digits, ms))
    return n.value, dict(zip(E_STAGES, list(ms)))


def e_digits(k, digits):
    """The reference's e recipe (test.cpp:12-19): returns (digit string,
    {stage: ms})."""
    buf = ctypes.create_string_buffer(digits + 64)
    n, st = e_digits_into(buf, k, digits)
    return buf.raw[:n].decode(), st


def threshold_words():
    return int(lib().nnh_threshold_words())


# ---- host-side helpers (no arithmetic on the product path) -----------------
def to_limbs(x, n=None):
    if n is None:
        n = max(1, (x.bit_length() + 63) // 64)
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype=np.uint64).copy()


def from_limbs(w):
    return int.from_bytes(np.ascontiguousarray(w, dtype=np.uint64).tobytes(), "little")


_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, top_nonzero=True):
    """n words of Vigna's splitmix64 stream from `seed` (state += gamma, then
    the two xor-shift-multiply rounds); the top word is forced nonzero."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    if top_nonzero and n and z[-1] == 0:
        z[-1] = 1
    return z
