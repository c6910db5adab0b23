"""The host copy team of the pinned staging ring (nnh_copyteam.hpp) on the
CPU under ThreadSanitizer: split copies of awkward sizes and offsets, every
byte checked, with the team stopped and restarted between rounds."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_copy_team_under_tsan(tmp_path):
    exe = str(tmp_path / "copyteam_check")
    src = os.path.join(ROOT, "tests", "cpp", "copyteam_check.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread", "-Wall", "-Wextra", src,
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout[-2000:] + r.stderr[-4000:]
