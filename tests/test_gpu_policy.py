"""The cost-model policy as the library's dispatch (nnh_policy_*).  A network
whose weights are zero predicts its output biases for every request, so a
policy with fixed arm preferences can be installed to drive every path the
dispatch can take: single products (nnh_mul, NNH_ALGO_AUTO/MODEL), batches
whose requests are planned per arm (chunked pieces joining the buckets and
sharing the shorter operand's transform, the reference split as two
pieces, schoolbook), with and without a graph, and the natural layer.  Every
product is compared with the oracle; the shipped policy (if any) is
reinstalled afterwards."""
import os

import numpy as np
import pytest

from conftest import rand_limbs
from test_gpu_batch import DevBatch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIPPED = os.path.join(ROOT, "natural_number_amd", "policy_gfx950.txt")

# predicted log2 ms per arm: school, karatsuba, ntt, ntt_split, ntt_chunk, ntt_chunk+1
PREF_CHUNK = [9, 9, 5, 6, 1, 2]
PREF_CHUNK1 = [9, 9, 5, 6, 2, 1]
PREF_SPLIT = [9, 9, 5, 1, 6, 7]
PREF_KARA = [2, 1, 5, 6, 7, 8]


def say(*a):
    print(*a, flush=True)


def constant_policy(cm, pref):
    m = cm.CostModel(seed=3)
    p = np.zeros(m.n_params, dtype=np.float32)
    p[-cm.N_ARMS:] = pref  # b3: the output biases
    m.set_params(p)
    return m


@pytest.fixture
def policy(nnh):
    from natural_number_amd import costmodel as cm

    models = []

    def install(pref):
        m = constant_policy(cm, pref)
        m.install()
        models.append(m)
        return m

    yield install
    for m in models:
        m.close()
    if os.path.exists(SHIPPED):
        nnh.policy_load(SHIPPED)
    else:
        nnh.policy_clear()


def uneven_ops(rng, n, big_hi=16.5):
    ops = []
    for i in range(n):
        big = int(2 ** rng.uniform(6, big_hi))
        small = max(16, int(big / 2 ** rng.uniform(0.3, 7)))
        a, b = rand_limbs(rng, big), rand_limbs(rng, small)
        ops.append((a, b) if i % 2 else (b, a))
    return ops


def test_policy_arm_queries(nnh, policy):
    policy(PREF_CHUNK)
    assert nnh.policy_active()
    arm, algo, param = nnh.policy_arm(100000, 300)
    assert (arm, algo) == (4, nnh.ALGO_NTT_CHUNK) and param == 10  # 2 x 300 words fit 2^10
    assert nnh.policy_arm(300, 100000)[0] == 4
    assert nnh.policy_arm(5000, 5000)[0] == 2  # chunked infeasible for a balanced product: one transform next
    assert nnh.policy_arm(10, 10)[0] == -1  # below the table: the reference's rule
    assert nnh.policy_arm(1 << 26, 1000)[0] == -1  # above the table
    nnh.policy_clear()
    assert not nnh.policy_active()
    assert nnh.policy_arm(100000, 300)[0] == -1


@pytest.mark.parametrize("pref", [PREF_CHUNK, PREF_CHUNK1, PREF_SPLIT, PREF_KARA],
                         ids=["chunk", "chunk+1", "split", "karatsuba"])
def test_policy_single_products(nnh, oracle, policy, pref):
    rng = np.random.default_rng(sum(pref))
    policy(pref)
    shapes = [(20000, 300), (300, 20000), (70000, 1000), (5000, 4999), (1 << 16, 1 << 12), (12345, 77), (400, 400)]
    for an, bn in shapes:
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        say(f"policy {pref}: {an} x {bn}: arm {nnh.policy_arm(an, bn)}")
        exp = oracle.mul(a, b)
        assert np.array_equal(nnh.mul(a, b), exp), (an, bn)
        assert np.array_equal(nnh.mul_algo(a, b, nnh.ALGO_MODEL), exp), (an, bn)
        assert np.array_equal(nnh.mul_algo(a, b, nnh.ALGO_REFERENCE), exp), (an, bn)


@pytest.mark.parametrize("pref", [PREF_CHUNK, PREF_SPLIT], ids=["chunk", "split"])
def test_policy_batches_vs_oracle(nnh, oracle, policy, pref):
    rng = np.random.default_rng(99 + pref[3])
    policy(pref)
    ops = uneven_ops(rng, 90)
    ops += [(rand_limbs(rng, 3000), rand_limbs(rng, 3000)), (rand_limbs(rng, 10), rand_limbs(rng, 9000))]
    # repeated B operands (shared transforms) next to chunked requests
    shared = rand_limbs(rng, 200)
    ops += [(rand_limbs(rng, int(2 ** rng.uniform(10, 15))), shared) for _ in range(6)]
    db = DevBatch(nnh, ops)
    n_chunked = sum(1 for a, b in ops if nnh.policy_arm(a.size, b.size, 64)[0] in (3, 4, 5))
    say(f"batch of {len(ops)} requests, {n_chunked} planned in pieces")
    assert n_chunked > 20
    for graph in (False, True):
        b = nnh.Batch(db.reqs, graph=graph)
        for rep in range(2 if graph else 1):
            db.clear()
            b.run()
            nnh.synchronize()
            for i, ((x, y), got) in enumerate(zip(ops, db.results())):
                assert np.array_equal(got, oracle.mul(x, y)), (i, x.size, y.size, graph, rep)
        b.close()
    # the same requests with the reference rule
    db.clear()
    b = nnh.Batch(db.reqs, graph=False, reference=True)
    b.run()
    nnh.synchronize()
    for (x, y), got in zip(ops, db.results()):
        assert np.array_equal(got, oracle.mul(x, y))
    b.close()


def test_policy_drives_natural_layer(nnh, oracle, policy):
    """Division, decimal output and e digits with a policy that sends
    unbalanced products to the chunked arm."""
    import hashlib
    import json

    from conftest import to_int

    rng = np.random.default_rng(4242)
    policy(PREF_CHUNK)
    a, b = rand_limbs(rng, 60000), rand_limbs(rng, 7000)
    q, r = nnh.divmod_(a, b)
    A, B = to_int(a), to_int(b)
    assert (to_int(q), to_int(r)) == divmod(A, B)
    x = rand_limbs(rng, 3000)
    assert nnh.to_dec(x) == str(to_int(x))
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "e_digits.json")))["e_k18_D1000000"]
    d, _ = nnh.e_digits(g["k"], g["digits"])
    assert hashlib.sha256(d.encode()).hexdigest() == g["sha256"]


def test_policy_file_round_trip(nnh, policy, tmp_path):
    from natural_number_amd import costmodel as cm

    m = policy(PREF_SPLIT)
    path = tmp_path / "p.txt"
    cm.save_policy(m, str(path))
    nnh.policy_clear()
    nnh.policy_load(str(path))
    assert nnh.policy_arm(100000, 300)[0] == 3
    bad = tmp_path / "bad.txt"
    bad.write_text("nnh-policy 1 3 6 5\n1 2 3\n")
    with pytest.raises(nnh.NNHError, match="malformed"):
        nnh.policy_load(str(bad))
    assert nnh.policy_arm(100000, 300)[0] == 3  # a failed load keeps the installed policy
