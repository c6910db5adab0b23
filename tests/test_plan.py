"""The host-only planner (natural_number_amd/csrc/nnh_plan.cpp) on the CPU,
built under AddressSanitizer and UndefinedBehaviorSanitizer: batch plans of
the full 4096-request uneven mix (the bench's batched workload) under the
reference rule and under policies that pick the chunked, split and
schoolbook arms, with every invariant of the plan checked (job order,
buckets, shared transforms, tiles, disjoint workspace regions, piece
windows tiling each chunked product); and small random plans executed on
the CPU as the device executes them, compared word for word with
schoolbook products.  The policy table's lookups and the arms' feasibility
are checked against the Python cost model's rules."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "natural_number_amd", "csrc")


@pytest.fixture(scope="module")
def plan_check(tmp_path_factory):
    d = tmp_path_factory.mktemp("plan")
    exe = str(d / "plan_check")
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-Wall", "-Wextra", "-I", CSRC,
           os.path.join(ROOT, "tests", "cpp", "plan_check.cpp"), os.path.join(CSRC, "nnh_plan.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe, d


def _run(args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run(args, capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-3000:])
    return r


def test_plans_of_the_uneven_mix_under_sanitizers(plan_check):
    from natural_number_amd.workloads import batch_mix_sizes

    exe, d = plan_check
    sizes = batch_mix_sizes()
    f = d / "mix.txt"
    f.write_text("".join(f"{a} {b}\n" for a, b in sizes))
    r = _run([exe, str(f)])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "plan_check ok" in r.stdout


def test_small_plans_executed_on_cpu_under_sanitizers(plan_check):
    exe, _ = plan_check
    r = _run([exe, "emulate"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "plan_check ok" in r.stdout


def test_arm_rules_match_the_python_cost_model():
    """nnh_arm_feasible / nnh_arm_call (the library's rules, no device
    needed) against costmodel.feasible / arm_call over random shapes."""
    from natural_number_amd import costmodel as cm

    rng = np.random.default_rng(5)
    for _ in range(3000):
        an = int(2 ** rng.uniform(0, 25))
        bn = int(2 ** rng.uniform(0, 25))
        assert np.array_equal(cm.library_feasible(an, bn), cm.feasible(an, bn)), (an, bn)
        for arm in range(cm.N_ARMS):
            assert cm.library_arm_call(arm, an, bn) == cm.arm_call(arm, an, bn), (arm, an, bn)
