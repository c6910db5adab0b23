"""What the trained cost model gains (config 5, reduced to run in seconds):
timings over a 10 x 10 grid of operand sizes up to 2^22 bits, the FP32 MLP
trained on the GPU by the library's kernels, and 300 held-out log-uniform
requests timed exhaustively (every feasible arm: the judge).  The trained
network must lower the regret of the untrained one, and its picks must not
cost more than the reference's fixed-threshold dispatch (NNH_ALGO_AUTO, the
method without a model) beyond timing noise; the gain is printed.  Every arm
is exact, so only time is at stake (the arms' exactness is tested against
the oracle in test_gpu_algos.py)."""
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_trained_model_gain_over_reference_policy(nnh):
    from natural_number_amd import costmodel as cm
    from natural_number_amd.workloads import log_uniform_requests

    bench = cm.Bench(max_words=1 << 16)
    try:
        pts = cm.grid(n_side=10, lo_bits=10, hi_bits=22, counts=(1,))
        say(f"cost model: timing {len(pts)} grid points")
        X, T, M = cm.collect(bench, pts)
        untrained = cm.CostModel(seed=7)
        model = cm.CostModel(seed=7)
        for lr in (3e-3, 1e-3, 3e-4):
            loss = model.train(X, T, M, steps=4000, lr=lr)
        reqs = log_uniform_requests(300, 0x4E4E2E1972800005, 10, 22)
        say("cost model: held-out evaluation, trained and untrained")
        ev = cm.evaluate(bench, model, reqs)
        ev0 = cm.evaluate(bench, untrained, reqs)
        untrained.close()
        model.close()
    finally:
        bench.close()
    gain = ev["total_ms_reference_policy"] / ev["total_ms_model"]
    say(f"train loss {loss:.4f}; regret trained {ev['regret_total']:.4f}, untrained {ev0['regret_total']:.4f}, "
        f"reference policy {ev['reference_policy_regret_total']:.4f}; gain over the reference policy {gain:.4f}")
    assert ev["regret_total"] < ev0["regret_total"]
    assert ev["total_ms_model"] <= 1.05 * ev["total_ms_reference_policy"]
