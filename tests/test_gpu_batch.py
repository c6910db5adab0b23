"""Batched path (bucketed persistent kernels, optional hipGraph) against the
oracle and against the single-product path.  Integer work: bit-exact."""
import numpy as np
import pytest

from conftest import rand_limbs, to_int

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class DevBatch:
    """Operands packed into one device buffer, outputs into another (device
    memory through the library's own runtime)."""

    def __init__(self, nnh, ops):
        sizes = [(a.size, b.size) for a, b in ops]
        host = np.concatenate([np.concatenate([a, b]) for a, b in ops])
        n_out = sum(an + bn for an, bn in sizes)
        self.inp = nnh.DeviceWords(host.size, host)
        self.out = nnh.DeviceWords(n_out)
        self.zeros = np.zeros(n_out, dtype=np.uint64)
        self.reqs, self.spans = [], []
        ia = oa = 0
        for (a, b), (an, bn) in zip(ops, sizes):
            da, db = self.inp.at(ia), self.inp.at(ia + an)
            if b is a:
                db = da
            self.reqs.append((self.out.at(oa), da, an, db, bn))
            self.spans.append((oa, an + bn))
            ia += an + bn
            oa += an + bn

    def clear(self):
        self.out.upload(self.zeros)

    def results(self):
        o = self.out.download()
        return [o[s:s + n] for s, n in self.spans]


def make_ops(rng, sizes, square_every=0):
    ops = []
    for i, (an, bn) in enumerate(sizes):
        a = rand_limbs(rng, an)
        if square_every and i % square_every == 0:
            ops.append((a, a))
        else:
            ops.append((a, rand_limbs(rng, bn)))
    return ops


def test_batch_mixed_vs_oracle(nnh, oracle):
    rng = np.random.default_rng(77)
    sizes = []
    for _ in range(160):
        an = int(2 ** rng.uniform(0, 14.5))
        bn = int(2 ** rng.uniform(0, 14.5))
        sizes.append((an, bn))
    sizes += [(384, 384), (383, 9000), (4096, 4096), (4097, 100), (30000, 20000)]
    ops = make_ops(rng, sizes, square_every=9)
    db = DevBatch(nnh, ops)
    for graph in (False, True):
        say(f"batch of {len(ops)} graph={graph}")
        db.clear()
        b = nnh.Batch(db.reqs, graph=graph)
        for rep in range(3 if graph else 1):
            if rep:
                db.clear()
            b.run()
            nnh.synchronize()
            for (a, bb), got in zip(ops, db.results()):
                exp = oracle.sqr(a) if bb is a else oracle.mul(a, bb)
                assert np.array_equal(got, exp), (a.size, bb.size, graph, rep)
        b.close()


def test_batch_config3_vs_oracle(nnh, oracle):
    """All 4096 products of the uneven mix, batched under a hipGraph with the
    installed policy's arms and with the reference's rule, and through one
    call per product: every product of each compared word for word with the
    oracle (16 host threads; the oracle's twiddle tables grown first)."""
    import ctypes
    from concurrent.futures import ThreadPoolExecutor

    from natural_number_amd.workloads import batch_mix_operands, batch_mix_sizes

    sizes = batch_mix_sizes()
    ops = list(batch_mix_operands(sizes))
    say(f"config3 batch: {len(ops)} products, {sum(a + b for a, b in sizes)} words, policy {nnh.policy_active()}")
    db = DevBatch(nnh, ops)
    results = {}
    for name, ref in (("policy", False), ("reference rule", True)):
        db.clear()
        b = nnh.Batch(db.reqs, graph=True, reference=ref)
        b.run()
        nnh.synchronize()
        results[name] = [r.copy() for r in db.results()]
        b.close()
    db.clear()
    for r in db.reqs:
        nnh.mul_device_async(*r, None)
    nnh.synchronize()
    results["single"] = db.results()
    smax = max((an + bn - 2).bit_length() for an, bn in sizes)
    tw = np.empty(1 << (smax - 1), dtype=np.uint64)
    oracle.L.nno_twiddles(tw.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 0, smax, 0)

    def check(i):
        exp = oracle.mul(*ops[i])
        return [k for k, v in results.items() if not np.array_equal(v[i], exp)]

    with ThreadPoolExecutor(16) as ex:
        bad = [(i, b) for i, b in enumerate(ex.map(check, range(len(ops)))) if b]
    assert not bad, bad[:10]


def test_batch_shared_b_transform(nnh, oracle):
    """Requests that repeat a B operand (same device pointer and size) share
    one transform of it inside the batch; products stay bit-exact.  Covers
    chunk-kernel-only scales, scales with top passes, a squaring request on
    the shared pointer, and B reused as another request's A."""
    rng = np.random.default_rng(91)
    groups = [(700, [600, 700, 650, 1, 699]), (3000, [2500, 3000, 2900, 3000]),
              (40000, [39000, 40000, 20000, 40000, 33333])]
    host, offs = [], []
    pos = 0

    def put(x):
        nonlocal pos
        host.append(x)
        offs.append(pos)
        pos += x.size
        return offs[-1]

    plan = []
    for bn, ans in groups:
        b = rand_limbs(rng, bn)
        ob = put(b)
        for an in ans:
            a = rand_limbs(rng, an)
            plan.append((put(a), a, ob, b))
        plan.append((ob, b, ob, b))          # square on the shared pointer
        c = rand_limbs(rng, bn // 2 + 1)
        plan.append((ob, b, put(c), c))      # the shared B as an A operand
    inp = nnh.DeviceWords(pos, np.concatenate(host))
    n_out = sum(a.size + b.size for _, a, _, b in plan)
    out = nnh.DeviceWords(n_out)
    reqs, spans, o = [], [], 0
    for oa, a, ob, b in plan:
        reqs.append((out.at(o), inp.at(oa), a.size, inp.at(ob), b.size))
        spans.append((o, a.size + b.size))
        o += a.size + b.size
    zeros = np.zeros(n_out, dtype=np.uint64)
    for graph in (False, True):
        say(f"shared-B batch of {len(reqs)} graph={graph}")
        bt = nnh.Batch(reqs, graph=graph)
        for rep in range(2):
            out.upload(zeros)
            bt.run()
            nnh.synchronize()
            got = out.download()
            for (oa, a, ob, b), (s, n) in zip(plan, spans):
                exp = oracle.sqr(a) if oa == ob else oracle.mul(a, b)
                assert np.array_equal(got[s:s + n], exp), (a.size, b.size, graph, rep)
        bt.close()
