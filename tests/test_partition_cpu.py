"""The device BVH builds place each element of a split segment directly where the
reference's in-place median partition (bvh.rs:243-256) leaves it (fgrain_amd/csrc/
partition.h).  Built for the host and run against the loop itself on random segments."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_partition_closed_form_equals_reference_loop(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "partition_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "fgrain_amd", "csrc"), "-o", str(exe),
                    os.path.join(ROOT, "tests", "cpu_hip", "partition_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n, bad = (int(x) for x in r.stdout.split()[1::2])
    assert n > 150000 and bad == 0
