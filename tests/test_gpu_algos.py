"""Every explicit algorithm arm is exact (against the oracle), including the
Karatsuba level's 65-bit coefficient carries and many-piece chunked
transforms; and the cost model's GPU gradients against float64 finite
differences."""
import numpy as np
import pytest

from conftest import rand_limbs

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (2, 2), (5, 3), (100, 100), (383, 1000), (1000, 1000), (3000, 50), (5000, 4000),
         (70000, 3000), (40000, 40001)]


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("algo", [1, 2, 3, 4, 5])
def test_arms_exact(nnh, oracle, algo):
    rng = np.random.default_rng(algo)
    for an, bn in SIZES:
        a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
        say(f"algo {nnh.ALGO_NAMES[algo]} {an}x{bn}")
        assert np.array_equal(nnh.mul_algo(a, b, algo), oracle.mul(a, b)), (an, bn)


def test_karatsuba_all_ones(nnh, oracle):
    for n, m in [(4096, 4096), (3001, 1999), (2, 2)]:
        a = np.full(n, 2**64 - 1, dtype=np.uint64)
        b = np.full(m, 2**64 - 1, dtype=np.uint64)
        say(f"karatsuba all-ones {n}x{m}")
        assert np.array_equal(nnh.mul_algo(a, b, nnh.ALGO_KARATSUBA), oracle.mul(a, b))


@pytest.mark.parametrize("sp", [9, 10, 12, 13])
def test_chunked_piece_scales(nnh, oracle, sp):
    rng = np.random.default_rng(sp)
    a, b = rand_limbs(rng, 100000), rand_limbs(rng, 300)
    say(f"chunked sp={sp}")
    assert np.array_equal(nnh.mul_algo(a, b, nnh.ALGO_NTT_CHUNK, sp), oracle.mul(a, b))


# (an, bn, sp): pieces of P = 2^sp - bn + 1 words grouped into G =
# ceil((P + bn) / P) buffers: G = 2, 3, 4 (schoolbook pieces for bn < 384,
# transform pieces sharing b's transform above), a last piece shorter than
# b, and G > 4 (the sequential form)
CHUNK_SHAPES = [(100000, 300, 10), (100000, 300, 9), (60000, 350, 9), (70001, 1000, 11), (80000, 1200, 11),
                (80000, 1500, 11), (1 << 20, 5000, 14), (50000, 400, 9), (40000, 1800, 11)]


@pytest.mark.parametrize("an,bn,sp", CHUNK_SHAPES)
def test_chunked_groups(nnh, oracle, an, bn, sp):
    rng = np.random.default_rng(an + bn + sp)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    P = (1 << sp) - bn + 1
    say(f"chunked {an}x{bn} sp={sp} G={-(-(P + bn) // P)}")
    want = oracle.mul(a, b)
    assert np.array_equal(nnh.mul_algo(a, b, nnh.ALGO_NTT_CHUNK, sp), want)
    assert np.array_equal(nnh.mul_algo(b, a, nnh.ALGO_NTT_CHUNK, sp), want)   # operands swapped


def test_mlp_gradients_vs_finite_differences(nnh):
    from natural_number_amd import costmodel as cm

    rng = np.random.default_rng(0)
    model = cm.CostModel(seed=3)
    n = 48
    X = rng.uniform(0.3, 1.0, size=(n, cm.N_IN)).astype(np.float32)
    T = rng.normal(0, 2, size=(n, cm.N_ARMS)).astype(np.float32)
    M = (rng.uniform(size=(n, cm.N_ARMS)) < 0.7).astype(np.float32)
    g, loss = model.gradients(X, T, M)
    p = model.params().astype(np.float64)
    Xd, Td, Md = X.astype(np.float64), T.astype(np.float64), M.astype(np.float64)
    assert abs(loss - cm.masked_mse(p, Xd, Td, Md)) < 1e-4 * max(1.0, loss)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-5
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (cm.masked_mse(pp, Xd, Td, Md) - cm.masked_mse(pm, Xd, Td, Md)) / (2 * h)
        assert abs(g[i] - fd) <= 2e-3 * max(1.0, abs(fd)), (i, g[i], fd)
    # training lowers the loss
    l1 = model.train(X, T, M, steps=300, lr=3e-3)
    assert l1 < loss
    model.close()


def test_butterfly_peak_rates(nnh):
    """The butterfly-issue ceiling runs for every prime and gives a plausible
    rate (between 10^10 and 10^14 butterflies per second); bad arguments fail."""
    for pr in range(3):
        r = nnh.butterfly_peak(pr, 32)
        assert 1e10 < r < 1e14, (pr, r)
    with pytest.raises(nnh.NNHError):
        nnh.butterfly_peak(3, 32)


def test_chunked_products_queued_back_to_back(nnh, oracle):
    # Chunked products plan their pieces as internal batches that are queued
    # without waiting: each one's job tables go up on the stream from one
    # pinned staging buffer, which the next call rewrites (and grows, for
    # the larger tables here) once the previous upload is done.  Several
    # such products queued with no synchronisation between them must each
    # match the oracle.
    shapes = [(20000, 1000, 11), (400000, 300, 10), (100000, 300, 10), (200000, 1500, 11), (60000, 350, 9)]
    rng = np.random.default_rng(77)
    ops = [(rand_limbs(rng, an), rand_limbs(rng, bn), sp) for an, bn, sp in shapes]
    st = nnh.Stream()
    bufs = []
    for a, b, sp in ops:
        da, db, dout = nnh.DeviceWords(a.size, a), nnh.DeviceWords(b.size, b), nnh.DeviceWords(a.size + b.size)
        bufs.append((da, db, dout))
    nnh.synchronize()
    for (a, b, sp), (da, db, dout) in zip(ops, bufs):
        nnh.mul_algo_device_async(dout.ptr, da.ptr, a.size, db.ptr, b.size, nnh.ALGO_NTT_CHUNK, sp, st.ptr)
    st.synchronize()
    for (a, b, sp), (da, db, dout) in zip(ops, bufs):
        say(f"queued chunked {a.size}x{b.size} sp={sp}")
        assert np.array_equal(dout.download(), oracle.mul(a, b)), (a.size, b.size, sp)
    for t in bufs:
        for x in t:
            x.free()
    st.close()


def test_chunked_arm_refuses_graph_capture(nnh, oracle):
    # The chunked arm's pieces go through an internal batch whose job tables
    # are uploaded from a staging buffer that later calls rewrite: under
    # stream capture the call is refused before anything is enqueued, and
    # the (empty) capture ends normally.
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    an, bn, sp = 100000, 1000, 11  # G = 2 groups of pieces: the batched form
    rng = np.random.default_rng(5)
    a, b = rand_limbs(rng, an), rand_limbs(rng, bn)
    da, db, dout = nnh.DeviceWords(an, a), nnh.DeviceWords(bn, b), nnh.DeviceWords(an + bn)
    st = nnh.Stream()
    nnh.synchronize()
    assert hip.hipStreamBeginCapture(ctypes.c_void_p(st.ptr), 2) == 0  # hipStreamCaptureModeRelaxed
    rc = nnh.lib().nnh_mul_algo_device_async(dout.ptr, da.ptr, an, db.ptr, bn, nnh.ALGO_NTT_CHUNK, sp, st.ptr)
    graph = ctypes.c_void_p()
    assert hip.hipStreamEndCapture(ctypes.c_void_p(st.ptr), ctypes.byref(graph)) == 0
    if graph.value:
        hip.hipGraphDestroy(graph)
    assert rc == 1, rc  # NNH_ERR_ARG
    assert "captured" in nnh.lib().nnh_last_error().decode()
    # the stream and the library still work
    nnh.mul_algo_device_async(dout.ptr, da.ptr, an, db.ptr, bn, nnh.ALGO_NTT_CHUNK, sp, st.ptr)
    st.synchronize()
    got = dout.download()
    st.close()
    for x in (da, db, dout):
        x.free()
    assert np.array_equal(got, oracle.mul(a, b))
