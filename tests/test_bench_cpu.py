"""bench.py's CPU baseline for the batched mix: the oracle's products over
the first requests of the mix, compared with reference outputs (here Python
integer products standing in for the GPU's); a changed word is counted.
Runs without a GPU."""
import os
import sys

import numpy as np
from conftest import to_int

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_batch_cpu_baseline_counts_identical_products():
    import bench
    from natural_number_amd.workloads import batch_mix_operands, batch_mix_sizes

    sizes = batch_mix_sizes()[:24]
    ops = list(batch_mix_operands(sizes))
    outs = [np.frombuffer((to_int(a) * to_int(b)).to_bytes(8 * (a.size + b.size), "little"), dtype=np.uint64)
            for a, b in ops]
    ref = np.concatenate(outs)
    r = bench.batch_cpu_baseline(ops, sizes, ref, max_threads=4)
    assert r["gpu_products_identical"] == len(ops)
    assert r["one_thread_products_per_s"] > 0 and r["all_cores_threads"] >= 1
    ref[outs[0].size + outs[1].size + 5] ^= 1 << 40   # one word of the third product
    assert bench.batch_cpu_baseline(ops, sizes, ref, max_threads=4)["gpu_products_identical"] == len(ops) - 1
