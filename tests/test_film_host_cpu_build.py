"""fg_film_optics and fg_film_composite -- film_host.cpp's host code (DevBuf growth and
reuse, the tile-table and Hann-ramp upload, the launch sequence) together with the
film.hip kernels -- built for the host against the HIP shim (tests/cpu_hip/) under
AddressSanitizer and UndefinedBehaviorSanitizer, with the compositing switched on
(FG_FILM_COMPOSITE=1), every device buffer an exact-size heap block and dynamic LDS
without slack.  One engine runs the GPU film tests' sequence of calls (optics on several
sizes, then the composite cases, including the 260 x 190 case whose first GPU launch
faulted) and every result is checked against the oracle."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from fgrain_amd import film as FM
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fgrain_amd", "csrc")
SHIM = os.path.join(ROOT, "tests", "cpu_hip")


def _scene(w, h, seed, letterbox=True):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    img = 40 + 160 * (xx / max(w - 1, 1)) * (0.5 + 0.5 * yy / max(h - 1, 1))
    for _ in range(4):
        cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(4, max(w, h, 20) / 4)
        img = np.where((xx - cx) ** 2 + (yy - cy) ** 2 < r * r, rng.uniform(0, 255), img)
    img = np.clip(np.round(img + rng.normal(0, 6, img.shape)), 0, 255).astype(np.uint8)
    if letterbox:
        b = max(h // 10, 1)
        img[:b] = 0
        img[h - b:] = 0
    return img


def _tiles(h_eff, step, pad, seed):
    rng = np.random.default_rng(seed)
    out = []
    for xs, ys, tw, th in O.film_tile_spans(h_eff.shape[1], h_eff.shape[0], step, pad):
        e = h_eff[ys:ys + th, xs:xs + tw]
        out.append((0.05 + 0.7 * e + 0.1 * rng.standard_normal(e.shape)).astype(np.float32).reshape(-1))
    return np.concatenate(out)


OPTICS = [(97, 61, True), (300, 200, True), (257, 130, False), (1, 1, False)]
COMPOSITE = [
    dict(w=260, h=190, opts=dict(grain_size_px=0.5, tile_size=64)),            # 5 x 3 tiles, radius 4
    dict(w=181, h=97, opts=dict(grain_size_px=0.3, tile_size=48, scanner_otf=False)),
    dict(w=530, h=120, opts=dict(grain_size_px=15.0, tile_size=120)),          # radius 48
]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("film_host_cpu")
    src = open(os.path.join(CSRC, "film.hip")).read()
    src = src.replace("    extern __shared__ float col[];", "    float *col = reinterpret_cast<float *>(cpu_dyn_lds);")
    assert "cpu_dyn_lds" in src
    (d / "film_cpu.cpp").write_text(src)
    exe = d / "film_host_driver"
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-ffp-contract=off", "-pthread", "-DFG_FILM_COMPOSITE=1",
           "-I", SHIM, "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-o", str(exe),
           os.path.join(SHIM, "film_host_driver.cpp"), os.path.join(CSRC, "film_host.cpp"), str(d / "film_cpu.cpp")]
    subprocess.run(cmd, check=True)
    return exe


def test_film_host_code_and_kernels_under_sanitizers(driver, tmp_path):
    jobs, want = [], []
    opts0 = FM.FilmRenderOptions()
    for w, h, lb in OPTICS:
        img = _scene(w, h, 5, lb)
        jobs.append((0, w, h, opts0, img, None, None))
        want.append(("optics", O.film_optics(img)[0]))
    for c in COMPOSITE:
        w, h = c["w"], c["h"]
        o = FM.FilmRenderOptions(**c["opts"])
        p = FM.film_plan(o, w, h)
        img = _scene(w, h, w + h)
        heff, _ = O.film_optics(img)
        tiles = _tiles(heff, p["tile_step"], p["pad"], 3)
        jobs.append((1, w, h, o, img, heff, tiles))
        ref = O.film_composite(img, heff, tiles, tile_step=p["tile_step"], pad=p["pad"],
                               grain_size_px=p["grain_size_px"], strength=p["strength"], preset=o.preset,
                               push=o.push, scanner_otf=o.scanner_otf, fold=1, intermediates=True)
        want.append(("composite", ref))
    with open(tmp_path / "jobs.bin", "wb") as f:
        for kind, w, h, o, img, heff, tiles in jobs:
            f.write(np.array([kind, w, h], np.uint32).tobytes())
            f.write(bytes(o.c()))
            f.write(np.ascontiguousarray(img, np.uint8).tobytes())
            if kind == 1:
                f.write(np.ascontiguousarray(heff, np.float32).tobytes())
                f.write(np.array([tiles.size], np.int64).tobytes())
                f.write(tiles.tobytes())
    r = subprocess.run([str(driver), str(tmp_path / "jobs.bin"), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    raw = open(tmp_path / "out.bin", "rb").read()
    at = 0
    bits = lambda a: np.ascontiguousarray(a, np.float32).view(np.uint32)
    for (kind, w, h, *_), (name, ref) in zip(jobs, want):
        n = w * h
        if kind == 0:
            heff = np.frombuffer(raw[at:at + 4 * n], np.float32).reshape(h, w)
            at += 4 * n
            assert (bits(heff) == bits(ref)).all(), (w, h)
            continue
        out = np.frombuffer(raw[at:at + n], np.uint8).reshape(h, w)
        at += n
        dn, dh, gr = (np.frombuffer(raw[at + 4 * n * k:at + 4 * n * (k + 1)], np.float32).reshape(h, w)
                      for k in range(3))
        at += 12 * n
        fs = np.frombuffer(raw[at:at + 12], np.float32)
        at += 12
        want_out, st, b = ref
        assert (bits(dn) == bits(b["d_norm"])).all(), (w, h)
        assert (bits(dh) == bits(b["d_norm"] - b["d_low"])).all(), (w, h)
        assert fs[0] == pytest.approx(st["high_mean"], rel=1e-6, abs=1e-9)
        assert fs[1] == pytest.approx(st["high_std"], rel=1e-6)
        assert fs[2] == pytest.approx(st["norm_factor"], rel=1e-6)
        np.testing.assert_allclose(gr, b["scanned"], rtol=1e-6, atol=1e-9)
        d = np.abs(out.astype(np.int16) - want_out.astype(np.int16))
        assert d.max() <= 1 and (d > 0).sum() <= 2
    assert at == len(raw)
