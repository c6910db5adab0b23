"""Synthetic workloads of the benchmark configurations (inputs only).

All randomness comes from Vigna's splitmix64 so the CPU oracle, the tests
and the benchmark regenerate identical operands from seeds.
"""
import math

import numpy as np

from . import splitmix64

MASK64 = (1 << 64) - 1


class SplitMix64:
    """Scalar splitmix64 stream (state += gamma; two xor-shift-multiply rounds)."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return self.next() / 2.0**64


BATCH_SEED = 0x4E4E2E1972800003
BATCH_OPERAND_SEED = 0x4E4E000000000000


def batch_mix_sizes(n=4096, seed=BATCH_SEED, lo_bits=12, hi_bits=24, max_ratio_log2=6):
    """Operand sizes (words) of the uneven batch: even requests balanced,
    m = n = 2^(lo + (hi-lo) u) bits; odd requests n = clamp(m / 2^(6 u'),
    2^lo, 2^hi) bits with the larger side chosen by bit 0 of a raw draw.
    Draw order per request: u, then (odd) u' and the raw word."""
    g = SplitMix64(seed)
    out = []
    for i in range(n):
        u = g.uniform()
        m_bits = 2.0 ** (lo_bits + (hi_bits - lo_bits) * u)
        if i % 2 == 0:
            n_bits = m_bits
        else:
            up = g.uniform()
            n_bits = min(max(m_bits / 2.0 ** (max_ratio_log2 * up), 2.0**lo_bits), 2.0**hi_bits)
            if g.next() & 1:
                m_bits, n_bits = n_bits, m_bits
        out.append((int(math.ceil(m_bits / 64)), int(math.ceil(n_bits / 64))))
    return out


def batch_mix_operands(sizes, seed=BATCH_OPERAND_SEED):
    """Operand words of request i: splitmix64 seeds seed+2i (A) and seed+2i+1 (B)."""
    for i, (an, bn) in enumerate(sizes):
        yield splitmix64(seed + 2 * i, an), splitmix64(seed + 2 * i + 1, bn)


def log_uniform_requests(n, seed, lo_bits=10, hi_bits=30):
    """Independent log-uniform operand sizes (words), for the cost model."""
    g = SplitMix64(seed)
    out = []
    for _ in range(n):
        a = 2.0 ** (lo_bits + (hi_bits - lo_bits) * g.uniform())
        b = 2.0 ** (lo_bits + (hi_bits - lo_bits) * g.uniform())
        out.append((max(1, int(math.ceil(a / 64))), max(1, int(math.ceil(b / 64)))))
    return out
