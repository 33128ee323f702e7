"""The scan walks 16-byte grid nodes (ScanQNode, fgrain_amd/csrc/fg_layout.h) in place of
the 32-byte float nodes: the grid test must enter every node whose float x-y box holds
the ray's origin strictly inside (the float test of transmittance.rs's walk for a ray
along -z), and its narrow form must accept a leaf only when the float test does, so each
ray meets the same leaves in the same order.  Built for the host from the same header
(-ffp-contract=off: the kernels' bits) and run on points within a few ulps of box edges,
around boxes and far outside the tree.  The same driver with the margins cut must find
violations, which shows that the samples reach the rounding band the margins cover."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fgrain_amd", "csrc")
DRIVER = os.path.join(ROOT, "tests", "cpu_hip", "scan_grid_check.cpp")


def _build(inc, exe):
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-pthread", "-I", str(inc), "-o", str(exe),
                    DRIVER], check=True)


def _run(exe, samples, seed):
    r = subprocess.run([str(exe), str(samples), str(seed), "4"], capture_output=True, text=True, timeout=300)
    m = re.search(r"samples (\d+) inside (\d+) surely (\d+) band (\d+) violations (\d+)", r.stdout)
    assert m, r.stdout + r.stderr
    return r, [int(x) for x in m.groups()]


@pytest.fixture(scope="module")
def gxx():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")


def test_grid_nodes_enter_every_node_and_accept_only_inside_points(gxx, tmp_path):
    exe = tmp_path / "scan_grid_check"
    _build(CSRC, exe)
    r, (n, inside, surely, band, bad) = _run(exe, 20_000_000, 21)
    assert r.returncode == 0 and bad == 0, r.stdout
    # points inside and outside in quantity, and both kinds of accepted leaf point: the
    # grid alone and, in the rounding band that the samples aim at, the float box
    assert n == 20_000_000 and 0.1 * n < inside < 0.9 * n
    assert surely > 0.3 * inside and band > 0.3 * inside


@pytest.mark.parametrize("old,new", [("r.may_lo[a] = scan_qint(floorf(q[a] + 0.5f));",
                                      "r.may_lo[a] = scan_qint(floorf(q[a] - 0.5f));"),
                                     ("r.in_hi[a] = scan_qint(ceilf(q[a] + 0.5f));",
                                      "r.in_hi[a] = scan_qint(ceilf(q[a] - 0.5f));")])
def test_samples_find_violations_when_a_margin_is_cut(gxx, tmp_path, old, new):
    src = open(os.path.join(CSRC, "fg_layout.h")).read()
    assert src.count(old) == 1, old
    inc = tmp_path / "weak"
    inc.mkdir()
    (inc / "fg_layout.h").write_text(src.replace(old, new))
    exe = tmp_path / "scan_grid_weak"
    _build(inc, exe)
    r, (n, inside, surely, band, bad) = _run(exe, 20_000_000, 22)
    assert bad > 0 and r.returncode == 1, r.stdout
