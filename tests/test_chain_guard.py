"""The admission rule of single-launch triangular solves, on the host: at
most four other streams of a device with a single-launch solve in flight,
same-stream launches counted once, completed launches freed."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_chain_admission_rule(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "chain_guard_test.cpp")
    exe = tmp_path / "chain_guard_test"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", src, "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
