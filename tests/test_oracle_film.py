"""Pins of the oracle's full-frame compositing restatement (oracle/film.c,
image_pipeline.rs:133-891) against closed forms, invariants and the reference's own
pipeline tests (image_pipeline.rs:1235-1326)."""
import numpy as np
import pytest

from oracle import oracle as O


# ------------------------------------------------------------------- plan --------
@pytest.mark.parametrize("grain,tile,step,pad", [
    (2.2, 320, 320, 32),     # the defaults: 384 * 2.2 = 844.8 -> tile_size
    (0.05, 320, 48, 12),     # round(19.2) = 19 -> 48, 48 / 10 -> 12
    (0.5, 320, 192, 19),     # 192, 19
    (1.0, 1000, 384, 38),
    (5.0, 1000, 1000, 48),   # 100 -> 48
    (0.126, 320, 48, 12),    # round(48.384) = 48
])
def test_plan_tile_step_and_overlap(grain, tile, step, pad):
    p = O.film_plan(1000, 700, grain_size_px=grain, tile_size=tile)
    assert (p["tile_step"], p["pad"]) == (step, pad)
    assert p["tiles_x"] == -(-1000 // step) and p["tiles_y"] == -(-700 // step)


def test_plan_scaling():
    # auto scale at half the reference resolution: grain x0.5, strength x sqrt(0.5)
    p = O.film_plan(3000, 2000, grain_size_px=2.2, auto_scale=True, reference_res=6000)
    assert p["grain_size_px"] == pytest.approx(1.1)
    assert p["strength"] == pytest.approx(np.sqrt(0.5))
    # manual scale 2: grain x2, strength x sqrt(2); grain clamps to [0.05, 15]
    p = O.film_plan(100, 100, grain_size_px=2.2, scale=2.0)
    assert p["grain_size_px"] == pytest.approx(4.4) and p["strength"] == pytest.approx(np.sqrt(2))
    assert O.film_plan(100, 100, grain_size_px=40.0)["grain_size_px"] == 15.0
    assert O.film_plan(100, 100, grain_size_px=0.001)["grain_size_px"] == 0.05
    # the ratio clamps to [0.05, 5] and the reference resolution to >= 512
    p = O.film_plan(100, 50, grain_size_px=1.0, auto_scale=True, reference_res=100)
    assert p["grain_size_px"] == pytest.approx(100 / 512)
    with pytest.raises(ValueError):
        O.film_plan(100, 100, tile_size=40)


def test_tile_spans_cover_the_frame():
    for w, h, step, pad in [(80, 64, 48, 12), (1000, 700, 320, 32), (321, 100, 320, 32)]:
        cover = np.zeros((h, w), np.int32)
        for xs, ys, tw, th in O.film_tile_spans(w, h, step, pad):
            cover[ys:ys + th, xs:xs + tw] += 1
        assert cover.min() >= 1 and cover.max() <= 4


# ------------------------------------------------------------------ luma ---------
def test_luma_weights_and_rounding():
    rgb = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [10, 20, 30]]], np.uint8)
    out = O.film_luma(rgb, "tri-x")
    # 255 * 0.42 = 107.1 -> 107, 255 * 0.46 = 117.3 -> 117, 255 * 0.12 = 30.6 -> 31
    assert out.tolist() == [[107, 117, 31, 255, round(10 * 0.42 + 20 * 0.46 + 30 * 0.12)]]
    assert O.film_luma(rgb, "t-max")[0, 0] == round(255 * 0.33)
    assert O.film_luma(rgb, "unknown")[0, 1] == round(255 * 0.48)


# ---------------------------------------------------------------- optics ---------
def _psf(w, h, x, y):
    """closed form of the interior impulse response: 0.9 n(x)n(y) + 0.1 w(x)w(y)."""
    kn = np.array([0.15, 0.70, 0.15])
    kw = np.array([0.03, 0.12, 0.20, 0.30, 0.20, 0.12, 0.03])
    out = np.zeros((h, w))
    out[y - 1:y + 2, x - 1:x + 2] += 0.9 * np.outer(kn, kn)
    out[y - 3:y + 4, x - 3:x + 4] += 0.1 * np.outer(kw, kw)
    return out


def test_optics_impulse_response_and_gate():
    src = np.zeros((40, 50), np.uint8)
    src[5:35, 4:46] = 1          # the gate: rows 5..35, columns 4..46
    src[20, 25] = 255
    h, m = O.film_optics(src)
    assert m == (5, 35, 4, 46)
    base = np.where(src > 0, 1.0 / 255, 0.0)
    impulse = _psf(50, 40, 25, 20) * (254.0 / 255)
    np.testing.assert_allclose(h[12:28, 17:33], (impulse + 1.0 / 255)[12:28, 17:33], atol=2e-7)
    assert (h[:5] == 0).all() and (h[35:] == 0).all() and (h[:, :4] == 0).all() and (h[:, 46:] == 0).all()
    # renormalised taps at the gate edge keep a flat field flat
    np.testing.assert_allclose(h[5:35, 4:8], base[5:35, 4:8], rtol=1e-6)


def test_optics_flat_and_empty():
    h, m = O.film_optics(np.full((9, 11), 200, np.uint8))
    assert m == (0, 9, 0, 11)
    np.testing.assert_allclose(h, 200 / 255, rtol=1e-6)
    h, m = O.film_optics(np.zeros((6, 7), np.uint8))
    assert m == (6, 6, 7, 7) and (h == 0).all()


# ----------------------------------------------------------- compositing ---------
def _single_tile(field, **kw):
    """A frame no larger than one tile: the Hann weight is 1, so d_norm is the tile."""
    h, w = field.shape
    src = np.full((h, w), 128, np.uint8)
    heff = np.full((h, w), 128 / 255, np.float32)
    return O.film_composite(src, heff, field.astype(np.float32), tile_step=max(w, h, 48), pad=12,
                            grain_size_px=kw.pop("grain", 2.2), intermediates=True, **kw)


def test_hann_partition_with_per_tile_constants():
    w, h, step, pad = 130, 100, 48, 12
    spans = O.film_tile_spans(w, h, step, pad)
    nx = -(-w // step)
    vals = [1.0 + 0.25 * i for i in range(len(spans))]
    tiles = np.concatenate([np.full(tw * th, v, np.float32) for v, (_, _, tw, th) in zip(vals, spans)])
    src = np.full((h, w), 128, np.uint8)
    _, _, b = O.film_composite(src, np.full((h, w), 0.5, np.float32), tiles, tile_step=step, pad=pad,
                               grain_size_px=2.2, intermediates=True)
    dn = b["d_norm"]
    # interiors hold their tile's value exactly
    assert dn[0, 0] == vals[0] and dn[60, 60] == vals[nx + 1]
    # across a vertical seam: (c_a fall(k) + c_b rise(k)) / (fall + rise), fall = 0.5 (1 + cos(pi k / 2 pad))
    y = 5
    for x in range(step - pad, step + pad):
        k = x - (step - pad)
        fall = 0.5 * (1 + np.cos(np.pi * k / (2 * pad)))
        rise = 0.5 * (1 - np.cos(np.pi * k / (2 * pad)))
        want = (vals[0] * fall + vals[1] * rise) / (fall + rise)
        assert dn[y, x] == pytest.approx(want, rel=2e-6), x
    # a shared field seen by every tile comes back unchanged
    yy, xx = np.mgrid[0:h, 0:w]
    f = (0.3 + 0.001 * xx + 0.002 * yy).astype(np.float32)
    tiles = np.concatenate([f[ys:ys + th, xs:xs + tw].reshape(-1) for xs, ys, tw, th in spans])
    _, _, b = O.film_composite(src, np.full((h, w), 0.5, np.float32), tiles, tile_step=step, pad=pad,
                               grain_size_px=2.2, intermediates=True)
    np.testing.assert_allclose(b["d_norm"], f, rtol=2e-6)


def test_lowpass_gaussian_closed_forms():
    n = 64
    # impulse well inside: d_low = outer product of the normalised Gaussian (sigma 3.3, radius 9)
    f = np.full((n, n), 1.0, np.float32)
    f[30, 31] += 1.0
    _, st, b = _single_tile(f)
    assert st["filter_sigma"] == pytest.approx(3.3, rel=1e-6) and st["k_rad"] == 9
    i = np.arange(-9, 10)
    g = np.exp(-0.5 * (i / 3.3) ** 2)
    g /= g.sum()
    want = 1.0 + np.outer(g, g)
    np.testing.assert_allclose(b["d_low"][21:40, 22:41], want, rtol=2e-6)
    # renormalised at the borders: a flat field stays flat, a ramp stays a ramp inside
    _, _, b = _single_tile(np.full((n, n), 0.7, np.float32))
    np.testing.assert_allclose(b["d_low"], 0.7, rtol=2e-6)
    yy, xx = np.mgrid[0:n, 0:n]
    ramp = (0.2 + 0.01 * xx - 0.004 * yy).astype(np.float32)
    _, _, b = _single_tile(ramp)
    np.testing.assert_allclose(b["d_low"][9:-9, 9:-9], ramp[9:-9, 9:-9], rtol=2e-6, atol=1e-6)
    # radius clamps: sigma 1.5 -> 4 taps; sigma 22.5 -> 48
    assert _single_tile(f, grain=0.5)[1]["k_rad"] == 4
    assert _single_tile(f, grain=15.0)[1]["k_rad"] == 48


def test_statistics_folds():
    rng = np.random.default_rng(1)
    f = (1.0 + 0.1 * rng.standard_normal((96, 96))).astype(np.float32)
    _, s0, b = _single_tile(f, fold=0)
    _, s1, _ = _single_tile(f, fold=1)
    d = b["d_norm"].astype(np.float64) - b["d_low"].astype(np.float64)
    assert s1["high_mean"] == pytest.approx(d.mean(), abs=1e-8)
    assert s1["high_std"] == pytest.approx(d.std(), rel=1e-6)
    assert s0["high_mean"] == pytest.approx(d.mean(), abs=1e-6)
    assert s0["high_std"] == pytest.approx(d.std(), rel=1e-5)
    assert s1["norm_factor"] == pytest.approx(min(0.038 / d.std(), 2.2), rel=1e-6)
    # the 0.035 floor of the standard deviation
    _, s, _ = _single_tile(np.full((50, 50), 0.4, np.float32))
    assert s["high_std"] == pytest.approx(0.035) and s["norm_factor"] == pytest.approx(0.038 / 0.035)


def _checker_frame(v, a=0.1, n=48):
    """single tile, uniform input value v, d_norm a +-a checkerboard around 1: d_high is
    +-a up to the Gaussian's residue at the Nyquist frequency (below 1e-30)."""
    yy, xx = np.mgrid[0:n, 0:n]
    f = (1.0 + a * np.where((xx + yy) % 2 == 0, 1.0, -1.0)).astype(np.float32)
    src = np.full((n, n), v, np.uint8)
    heff, _ = O.film_optics(src)
    return O.film_composite(src, heff, f, tile_step=48, pad=12, grain_size_px=2.2, scanner_otf=False,
                            intermediates=True)


def _visibility(b, st):
    """grain / (sharpened high-pass x norm factor) over the interior (Hann weight 1)."""
    dd = (b["d_norm"].astype(np.float64) - b["d_low"]) - st["high_mean"]
    s = np.sign(dd) * np.abs(dd) ** 0.85 * st["norm_factor"]
    return (b["grain"] / s)[9:-9, 9:-9]


def test_gates_and_midtone_at_closed_form_points():
    a = 0.1
    for v, vis in [(0, 0.0), (1, 0.0),             # below the toe: no grain
                   (97, 1.0),                      # base^0.72 = 0.5 -> sin = 1 (0.3804 ~ 0.3816)
                   (255, 0.0)]:                    # above the shoulder: fused
        _, st, b = _checker_frame(v, a)
        g = _visibility(b, st)
        np.testing.assert_allclose(g, vis, atol=1e-4, err_msg=str(v))
    # smoothstep midpoints: base 0.0725 -> shadow gate 0.5, base 0.795 -> highlight 0.5
    for v, gate in [(0.0725 * 255, 0.5), (0.795 * 255, 0.5)]:
        v8 = int(round(v))
        base = v8 / 255
        mid = np.sin(np.pi * base ** 0.72)
        t = (base - 0.005) / 0.135
        u = (base - 0.65) / 0.29
        sg = 1.0 if t >= 1 else t * t * (3 - 2 * t)
        hg = 1.0 if u <= 0 else 1 - u * u * (3 - 2 * u)
        assert sg * hg == pytest.approx(gate, abs=0.03)
        _, st, b = _checker_frame(v8, a)
        np.testing.assert_allclose(_visibility(b, st), sg * hg * mid, rtol=1e-5)
    # no push: the tone is the base value
    _, _, b = _checker_frame(128, a)
    assert b["tone"][10, 10] == pytest.approx(128 / 255, rel=1e-6)


def test_push_s_curve_invariants():
    n = 48
    for v in (40, 100, 128, 155, 215):
        src = np.full((n, n), v, np.uint8)
        heff, _ = O.film_optics(src)
        f = np.ones((n, n), np.float32)
        kw = dict(tile_step=48, pad=12, grain_size_px=2.2, intermediates=True)
        t0 = O.film_composite(src, heff, f, **kw)[2]["tone"][5, 5]
        t1 = O.film_composite(src, heff, f, push=1.0, **kw)[2]["tone"][5, 5]
        b = v / 255
        if v == 128:
            assert abs(t1 - 0.5) < abs(t0 - 0.5) * 1.2 + 1e-3
        # steeper midtones: pushed tones move away from 0.5
        assert abs(t1 - 0.5) >= abs(t0 - 0.5) - 1e-6 or abs(b - 0.5) < 0.01
        # symmetric about 0.5 (gamma curve of 2b and 2(1-b))
        src2 = np.full((n, n), 255 - v, np.uint8)
        heff2, _ = O.film_optics(src2)
        t2 = O.film_composite(src2, heff2, f, push=1.0, **kw)[2]["tone"][5, 5]
        gamma = 1.06
        want = 0.5 * (2 * b) ** gamma if b <= 0.5 else 1 - 0.5 * (2 * (1 - b)) ** gamma
        assert t1 == pytest.approx(want, rel=1e-5)
        assert t1 + t2 == pytest.approx(1.0, abs=1e-5)


def test_otf_is_the_scanner_kernel():
    a = 0.1
    _, _, b0 = _checker_frame(97, a)
    yy, xx = np.mgrid[0:48, 0:48]
    # the OTF of a checkerboard: 0.93 - 2 * 0.035 per axis in the interior
    _, st, b = _checker_frame(97, a)
    src = np.full((48, 48), 97, np.uint8)
    heff, _ = O.film_optics(src)
    f = (1.0 + a * np.where((xx + yy) % 2 == 0, 1.0, -1.0)).astype(np.float32)
    _, _, b1 = O.film_composite(src, heff, f, tile_step=48, pad=12, grain_size_px=2.2, scanner_otf=True,
                                intermediates=True)
    # (outside the 10-pixel border where the low-pass sees the frame edge)
    np.testing.assert_allclose(b1["scanned"][10:-10, 10:-10], b0["grain"][10:-10, 10:-10] * (0.93 - 0.07) ** 2,
                               rtol=1e-5)
    np.testing.assert_allclose(b1["scanned"], O.otf3(b0["grain"]), rtol=1e-6, atol=1e-9)


# ------------------------------------------------- the reference's own tests ------
def _dneg_tiles(h_eff, step, pad, level=0.6, seed=0):
    """Stand-in D_neg per tile: a density that follows the exposure plus seeded grain noise."""
    h, w = h_eff.shape
    rng = np.random.default_rng(seed)
    out = []
    for xs, ys, tw, th in O.film_tile_spans(w, h, step, pad):
        e = h_eff[ys:ys + th, xs:xs + tw]
        out.append((0.05 + level * e + 0.08 * rng.standard_normal(e.shape)).astype(np.float32).reshape(-1))
    return np.concatenate(out)


@pytest.mark.parametrize("size", [2.4, 10.0, 20.0])
def test_letterbox_and_shadow_zero_variance(size):
    """image_pipeline.rs:1235-1277 on the compositing stage: letterbox rows and the
    silhouette core are exactly 0 whatever the grain."""
    img = np.full((64, 64), 200, np.uint8)
    img[:8] = 0
    img[56:] = 0
    img[24:40, 24:40] = 0
    p = O.film_plan(64, 64, grain_size_px=size, tile_size=64)
    heff, _ = O.film_optics(img)
    out, _ = O.film_composite(img, heff, _dneg_tiles(heff, p["tile_step"], p["pad"]),
                              tile_step=p["tile_step"], pad=p["pad"], grain_size_px=p["grain_size_px"])
    assert (out[:8] == 0).all() and (out[56:] == 0).all() and (out[26:38, 26:38] == 0).all()


def test_highlight_fusion_burnout():
    """image_pipeline.rs:1279-1326 on the compositing stage: the blown-out half fuses."""
    img = np.full((64, 64), 128, np.uint8)
    img[:, :32] = 255
    p = O.film_plan(64, 64, grain_size_px=2.4, tile_size=64)
    heff, _ = O.film_optics(img)
    out, _ = O.film_composite(img, heff, _dneg_tiles(heff, p["tile_step"], p["pad"]),
                              tile_step=p["tile_step"], pad=p["pad"], grain_size_px=p["grain_size_px"])
    fog = out[4:60, 4:28]
    assert fog.min() >= 253 and fog.max() == 255
    assert out[:, 40:].std() > 0.5   # the midtone half keeps its grain
