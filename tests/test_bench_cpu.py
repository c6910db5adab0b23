"""bench.py on the CPU: the batched mix's CPU baseline (the oracle's products
over the first requests of the mix, compared with reference outputs, here
Python integer products standing in for the GPU's; a changed word is
counted), and the refusal of more than one GPU.  Runs without a GPU."""
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


def test_bench_refuses_more_than_one_gpu():
    # one GPU per box: --gpus N > 1 exits with a message, before any GPU work
    import subprocess

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
