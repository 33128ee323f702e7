"""The film kernels' source (fgrain_amd/csrc/film.hip) built for the host against a
small HIP shim (tests/cpu_hip/cpu_hip.h) and compared with the oracle: the optics, Hann
blend and low-pass bit for bit, the grain and 8-bit output bit for bit as well (the host
build calls the same libm as the oracle), the statistics to f64 summation order.  This
pins the source of the compositing kernels on the CPU; their GPU runs are in
tests/test_gpu_film.py."""
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


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("film_cpu")
    src = open(os.path.join(CSRC, "film.hip")).read()
    src = src.replace("#include <hip/hip_runtime.h>", '#include "cpu_hip.h"')
    src = src.replace("    extern __shared__ float col[];", "    float *col = reinterpret_cast<float *>(cpu_dyn_lds);")
    assert "cpu_dyn_lds" in src
    (d / "film_cpu.cpp").write_text(src)
    hdr = open(os.path.join(CSRC, "film.h")).read().replace("#include <hip/hip_runtime.h>", '#include "cpu_hip.h"')
    (d / "film.h").write_text(hdr)
    shutil.copy(os.path.join(SHIM, "cpu_hip.h"), d / "cpu_hip.h")
    exe = d / "film_driver"
    subprocess.run(["g++", "-std=c++20", "-O1", "-ffp-contract=off", "-pthread", "-I", str(d), "-o", str(exe),
                    os.path.join(SHIM, "film_driver.cpp"), str(d / "film_cpu.cpp")], check=True)
    return exe


def _scene(w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    img = 30 + 170 * (xx / (w - 1)) * (0.5 + 0.5 * yy / (h - 1)) + rng.normal(0, 8, (h, w))
    img = np.clip(np.round(img), 0, 255).astype(np.uint8)
    img[: h // 8] = 0
    img[h // 2:h // 2 + 4, w // 3:w // 3 + 6] = 0
    return img


@pytest.mark.parametrize("w,h,opts", [
    (100, 70, dict(grain_size_px=0.5, tile_size=48)),                  # 3 x 2 tiles, radius 4
    (90, 61, dict(preset="tri-x-1600", grain_size_px=1.0, tile_size=48, scanner_otf=False)),  # push 2, no OTF
])
def test_cpu_build_of_film_kernels_matches_oracle(driver, tmp_path, w, h, opts):
    o = FM.FilmRenderOptions(**opts)
    p = FM.film_plan(o, w, h)
    img = _scene(w, h, w)
    rng = np.random.default_rng(1)
    heff_o, _ = O.film_optics(img)
    spans = O.film_tile_spans(w, h, p["tile_step"], p["pad"])
    tiles = np.concatenate([(0.05 + 0.7 * heff_o[ys:ys + th, xs:xs + tw] + 0.1 * rng.standard_normal((th, tw)))
                            .astype(np.float32).reshape(-1) for xs, ys, tw, th in spans])
    base_amp = np.float32(p["base_nominal"]) * (np.float32(1.0) + np.float32(0.15) * np.float32(p["push"]))
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(np.array([w, h, p["tile_step"], p["pad"], p["k_rad"], int(o.scanner_otf)], np.uint32).tobytes())
        f.write(np.array([p["strength"], base_amp, p["push"], p["filter_sigma"]], np.float32).tobytes())
        f.write(img.tobytes())
        f.write(tiles.tobytes())
    subprocess.run([str(driver), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], check=True, timeout=300)
    raw = open(tmp_path / "out.bin", "rb").read()
    n = w * h

    def f32(k):
        return np.frombuffer(raw[4 * k * n:4 * (k + 1) * n], np.float32).reshape(h, w)
    heff, dn, dh, grain = f32(0), f32(1), f32(2), f32(3)
    out = np.frombuffer(raw[16 * n:17 * n], np.uint8).reshape(h, w)
    fs = np.frombuffer(raw[17 * n:17 * n + 16], np.float32)
    want, st, b = O.film_composite(img, heff_o, tiles, tile_step=p["tile_step"], pad=p["pad"],
                                   grain_size_px=p["grain_size_px"], strength=p["strength"], preset=o.preset,
                                   push=o.push, scanner_otf=o.scanner_otf, fold=1, intermediates=True)
    bits = lambda a: np.ascontiguousarray(a, np.float32).view(np.uint32)
    assert (bits(heff) == bits(heff_o)).all()
    assert (bits(dn) == bits(b["d_norm"])).all()
    assert (bits(dh) == bits(b["d_norm"] - b["d_low"])).all()
    assert fs[0] == pytest.approx(st["high_mean"], rel=1e-6, abs=1e-9)
    assert fs[1] == pytest.approx(st["high_std"], rel=1e-6)
    np.testing.assert_allclose(grain, b["scanned"], rtol=1e-6, atol=1e-9)
    d = np.abs(out.astype(np.int16) - want.astype(np.int16))
    assert d.max() <= 1 and (d > 0).sum() <= 2
