"""The transport tests a leaf's crystals in two phases: a cheap conservative pre-test
(crystal_may_absorb, fgrain_amd/csrc/fg_layout.h) and then crystal_hit only on the
crystals it keeps, so the pre-test must never drop a crystal on which the float crystal
test records an event (transport.rs:109-131).  Built for the host from the same header
(-ffp-contract=off: the kernels' bits) and run on samples aimed at the pre-test's margins:
far origins, grazing rays, segment ends at the crystal, flat ellipsoids.  The same driver
with the margins cut a hundredfold must find violations, which shows that the samples
reach the cases the margins are there for."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fgrain_amd", "csrc")
DRIVER = os.path.join(ROOT, "tests", "cpu_hip", "absorb_pretest_check.cpp")


def _build(inc, exe):
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-pthread", "-I", str(inc), "-o", str(exe),
                    DRIVER], check=True)


def _run(exe, samples, seed):
    r = subprocess.run([str(exe), str(samples), str(seed), "4"], capture_output=True, text=True, timeout=300)
    m = re.search(r"samples (\d+) events (\d+) rejected (\d+) violations (\d+)", r.stdout)
    assert m, r.stdout + r.stderr
    return r, [int(x) for x in m.groups()]


@pytest.fixture(scope="module")
def gxx():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")


def test_pretest_never_drops_an_event(gxx, tmp_path):
    exe = tmp_path / "absorb_pretest_check"
    _build(CSRC, exe)
    r, (n, events, rejected, bad) = _run(exe, 40_000_000, 11)
    assert r.returncode == 0 and bad == 0, r.stdout
    # the samples hold events and rejections in quantity, not one kind only
    assert n > 39_000_000 and events > 0.2 * n and rejected > 0.3 * n


def test_samples_find_violations_when_the_margins_are_cut(gxx, tmp_path):
    src = open(os.path.join(CSRC, "fg_layout.h")).read()
    weak = src
    for old, new in [("const float m = 1e-5f * (pn + R);", "const float m = 1e-7f * (pn + R);"),
                     ("(1e-5f + 6e-6f * rho)", "(1e-7f + 6e-8f * rho)"),
                     ("hmax * (8e-3f * rho)", "hmax * (8e-5f * rho)")]:
        assert weak.count(old) == 1, old
        weak = weak.replace(old, new)
    inc = tmp_path / "weak"
    inc.mkdir()
    (inc / "fg_layout.h").write_text(weak)
    exe = tmp_path / "absorb_pretest_weak"
    _build(inc, exe)
    r, (n, events, rejected, bad) = _run(exe, 40_000_000, 12)
    assert bad > 0 and r.returncode == 1, r.stdout


def test_samples_find_violations_when_the_packed_record_rounds_the_wrong_way(gxx, tmp_path):
    # the frame transport's packed pre-test record must round the largest semi-axis up;
    # rounded down it rejects segments the float pre-test keeps
    src = open(os.path.join(CSRC, "fg_layout.h")).read()
    old = "f16_bits_dir(fmaxf(fmaxf(h[0], h[1]), h[2]), true)"
    assert src.count(old) == 1
    inc = tmp_path / "weak"
    inc.mkdir()
    (inc / "fg_layout.h").write_text(src.replace(old, old.replace("true)", "false)")))
    exe = tmp_path / "absorb_pretest_packed_weak"
    _build(inc, exe)
    r, (n, events, rejected, bad) = _run(exe, 4_000_000, 13)
    assert bad > 0 and r.returncode == 1 and "(packed record)" in r.stdout, r.stdout
