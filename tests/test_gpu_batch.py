"""Batched path (bucketed persistent kernels, optional hipGraph) against the
oracle and against the single-product path.  Integer work: bit-exact."""
import numpy as np
import pytest

from conftest import rand_limbs, to_int

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class DevBatch:
    """Operands packed into one device buffer, outputs into another."""

    def __init__(self, torch, ops):
        self.torch = torch
        sizes = [(a.size, b.size) for a, b in ops]
        n_in = sum(an + bn for an, bn in sizes)
        n_out = sum(an + bn for an, bn in sizes)
        host = np.concatenate([np.concatenate([a, b]) for a, b in ops]) if ops else np.zeros(0, np.uint64)
        self.inp = torch.from_numpy(host.view(np.int64)).cuda()
        self.out = torch.zeros(n_out, dtype=torch.int64, device="cuda")
        self.reqs, self.spans = [], []
        ip, op = self.inp.data_ptr(), self.out.data_ptr()
        ia = oa = 0
        for (a, b), (an, bn) in zip(ops, sizes):
            da, db = ip + 8 * ia, ip + 8 * (ia + an)
            if b is a:
                db, bn_ = da, an
            else:
                bn_ = bn
            self.reqs.append((op + 8 * oa, da, an, db, bn_))
            self.spans.append((oa, an + bn))
            ia += an + bn
            oa += an + bn

    def results(self):
        o = self.out.cpu().numpy().view(np.uint64)
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


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def test_batch_mixed_vs_oracle(nnh, oracle, torch):
    rng = np.random.default_rng(77)
    sizes = []
    for _ in range(160):
        an = int(2 ** rng.uniform(0, 14.5))
        bn = int(2 ** rng.uniform(0, 14.5))
        sizes.append((an, bn))
    sizes += [(384, 384), (383, 9000), (4096, 4096), (4097, 100), (30000, 20000)]
    ops = make_ops(rng, sizes, square_every=9)
    db = DevBatch(torch, ops)
    for graph in (False, True):
        say(f"batch of {len(ops)} graph={graph}")
        db.out.zero_()
        torch.cuda.synchronize()
        b = nnh.Batch(db.reqs, graph=graph)
        for rep in range(3 if graph else 1):
            if rep:
                db.out.zero_()
                torch.cuda.synchronize()
            b.run()
            torch.cuda.synchronize()
            for (a, bb), got in zip(ops, db.results()):
                exp = oracle.sqr(a) if bb is a else oracle.mul(a, bb)
                assert np.array_equal(got, exp), (a.size, bb.size, graph, rep)
        b.close()


def test_batch_config3_matches_single_path(nnh, torch):
    """All 4096 products of the uneven mix: batched (graph) == one call per product,
    and every product checked modulo two primes independently of the GPU."""
    from natural_number_amd.workloads import batch_mix_operands, batch_mix_sizes

    sizes = batch_mix_sizes()
    ops = list(batch_mix_operands(sizes))
    say(f"config3 batch: {len(ops)} products, {sum(a + b for a, b in sizes)} words")
    db = DevBatch(torch, ops)
    b = nnh.Batch(db.reqs, graph=True)
    b.run()
    torch.cuda.synchronize()
    batched = [r.copy() for r in db.results()]
    db.out.zero_()
    torch.cuda.synchronize()
    for r in db.reqs:
        nnh.mul_device_async(*r, None)
    torch.cuda.synchronize()
    single = db.results()
    for i, (x, y) in enumerate(zip(batched, single)):
        assert np.array_equal(x, y), i
    for q in (2**61 - 1, (1 << 64) - 59):
        for (a, bb), got in zip(ops[::7], batched[::7]):
            assert to_int(got) % q == (to_int(a) % q) * (to_int(bb) % q) % q
    b.close()
