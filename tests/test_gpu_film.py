"""Full-frame film render on the GPU (fg_render_film_bw and its stage entry points)
against the oracle's restatement of image_pipeline.rs:133-891 and the reference's own
pipeline tests (image_pipeline.rs:1235-1326)."""
import numpy as np
import pytest

import fgrain_amd as F
from fgrain_amd import film as FM
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# The compositing kernels (Hann blend, Gaussian separation, statistics, grain, output)
# are switched off in the library: their first launch on the MI355X ended in a device
# memory fault whose cause is not found yet (the kernels' source built for the CPU
# against a HIP shim matches the oracle bit for bit).  The tests that launch them stay
# here, skipped, until the cause is fixed.
COMPOSITE_OFF = ("compositing kernels switched off: device memory fault on their first GPU launch, "
                 "cause not found yet")


@pytest.fixture(scope="module")
def engine():
    return F.GpuEngine.new()


def _scene(w, h, seed, letterbox=True):
    """Seeded 8-bit scene: smooth gradient, a few bright and dark discs, black letterbox
    bars and a zero-valued silhouette."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    img = 40 + 160 * (xx / max(w - 1, 1)) * (0.5 + 0.5 * yy / max(h - 1, 1))
    for _ in range(4):
        cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(4, max(w, h, 20) / 4)
        img = np.where((xx - cx) ** 2 + (yy - cy) ** 2 < r * r, rng.uniform(0, 255), img)
    img = img + rng.normal(0, 6, img.shape)
    img = np.clip(np.round(img), 0, 255).astype(np.uint8)
    if letterbox:
        b = max(h // 10, 1)
        img[:b] = 0
        img[h - b:] = 0
        img[h // 2 - 3:h // 2 + 3, w // 3:w // 3 + 8] = 0
    return img


def _tiles(h_eff, step, pad, seed):
    rng = np.random.default_rng(seed)
    out = []
    for xs, ys, tw, th in O.film_tile_spans(h_eff.shape[1], h_eff.shape[0], step, pad):
        e = h_eff[ys:ys + th, xs:xs + tw]
        out.append((0.05 + 0.7 * e + 0.1 * rng.standard_normal(e.shape)).astype(np.float32).reshape(-1))
    return np.concatenate(out)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("w,h,letterbox", [(97, 61, True), (300, 200, True), (257, 130, False), (1, 1, False)])
def test_optics_bit_exact(engine, w, h, letterbox):
    img = _scene(w, h, 5, letterbox)
    got = FM.film_optics(img, engine)
    want, _ = O.film_optics(img)
    assert (_bits(got) == _bits(want)).all()


def test_optics_all_black_and_margins(engine):
    assert (FM.film_optics(np.zeros((40, 33), np.uint8), engine) == 0).all()
    img = np.zeros((50, 70), np.uint8)
    img[7:41, 9:60] = 180
    img[20, 30] = 3
    assert (_bits(FM.film_optics(img, engine)) == _bits(O.film_optics(img)[0])).all()


def test_luma_bit_exact(engine):
    rgb = np.random.default_rng(2).integers(0, 256, (45, 67, 3), dtype=np.uint8)
    for preset in ("tri-x", "hp5", "t-max", "something-else"):
        assert (FM.film_luma(rgb, preset, engine) == O.film_luma(rgb, preset)).all()


CASES = [
    dict(w=260, h=190, opts=dict(grain_size_px=0.5, tile_size=64)),                 # 5 x 3 tiles, radius 4
    dict(w=300, h=211, opts=dict(grain_size_px=2.2, tile_size=100)),                # radius 9, 3 x 3 tiles
    dict(w=530, h=120, opts=dict(grain_size_px=15.0, tile_size=120)),               # radius 48, > 256 wide
    dict(w=200, h=150, opts=dict(preset="tri-x-1600", grain_size_px=1.0, tile_size=64)),   # push 2: tone curve
    dict(w=181, h=97, opts=dict(grain_size_px=0.3, tile_size=48, scanner_otf=False)),       # no OTF
    dict(w=150, h=150, opts=dict(preset="t-max", grain_size_px=2.0, tile_size=320, grain_strength=0.6)),  # 1 tile
]


@pytest.mark.skip(reason=COMPOSITE_OFF)
@pytest.mark.parametrize("case", CASES, ids=[f"{c['w']}x{c['h']}" for c in CASES])
def test_composite_matches_oracle(engine, case):
    w, h = case["w"], case["h"]
    opts = FM.FilmRenderOptions(**case["opts"])
    plan = FM.film_plan(opts, w, h)
    img = _scene(w, h, w + h)
    heff, _ = O.film_optics(img)
    tiles = _tiles(heff, plan["tile_step"], plan["pad"], 3)
    out, st, b = FM.film_composite(img, heff, tiles, opts, engine, intermediates=True)
    want, ost, ob = O.film_composite(img, heff, tiles, tile_step=plan["tile_step"], pad=plan["pad"],
                                     grain_size_px=plan["grain_size_px"], strength=plan["strength"],
                                     preset=opts.preset, push=opts.push, scanner_otf=opts.scanner_otf,
                                     fold=1, intermediates=True)
    # blend and low-pass: the reference's f32 operations, bit for bit
    assert (_bits(b["d_norm"]) == _bits(ob["d_norm"])).all()
    assert (_bits(b["d_high"]) == _bits(ob["d_norm"] - ob["d_low"])).all()
    # statistics: f64 sums of the same terms
    assert st["high_mean"] == pytest.approx(ost["high_mean"], rel=1e-6, abs=1e-9)
    assert st["high_std"] == pytest.approx(ost["high_std"], rel=1e-6)
    assert st["norm_factor"] == pytest.approx(ost["norm_factor"], rel=1e-6)
    assert st["k_rad"] == ost["k_rad"] and st["filter_sigma"] == ost["filter_sigma"]
    # grain: the device's sinf / powf round differently in the last bits
    np.testing.assert_allclose(b["grain"], ob["scanned"], rtol=2e-5, atol=2e-7)
    d = np.abs(out.astype(np.int16) - want.astype(np.int16))
    assert d.max() <= 1 and (d > 0).mean() <= 1e-3, (d.max(), (d > 0).sum())


@pytest.mark.skip(reason=COMPOSITE_OFF)
def test_composite_reference_fold_is_close(engine):
    """The reference's sequential f32 statistics differ from the f64 sums only in the
    fold's own rounding at these sizes."""
    w, h = 300, 211
    opts = FM.FilmRenderOptions(tile_size=100)
    plan = FM.film_plan(opts, w, h)
    img = _scene(w, h, 9)
    heff, _ = O.film_optics(img)
    tiles = _tiles(heff, plan["tile_step"], plan["pad"], 4)
    out, st = FM.film_composite(img, heff, tiles, opts, engine)
    want, ost = O.film_composite(img, heff, tiles, tile_step=plan["tile_step"], pad=plan["pad"],
                                 grain_size_px=plan["grain_size_px"], fold=0)
    assert st["high_std"] == pytest.approx(ost["high_std"], rel=1e-4)
    d = np.abs(out.astype(np.int16) - want.astype(np.int16))
    assert d.max() <= 1 and (d > 0).mean() <= 2e-3


def test_composite_layout_errors(engine):
    img = np.zeros((40, 40), np.uint8)
    with pytest.raises(F.GpuError):
        FM.film_composite(img, np.zeros((40, 40), np.float32), np.zeros(10, np.float32),
                          FM.FilmRenderOptions(tile_size=48), engine)


@pytest.mark.skip(reason=COMPOSITE_OFF)
def test_film_render_equals_per_tile_renders(engine):
    """fg_render_film_bw = optics -> render_tile_all_gpu per overlapping tile (tile seeds,
    image_pipeline.rs:430-434) -> compositing, bit for bit; report counts are the tiles'."""
    w, h = 150, 110
    opts = FM.FilmRenderOptions(tile_size=64, seed=7)
    plan = FM.film_plan(opts, w, h)
    img = _scene(w, h, 11)
    rep, out = FM.render_film_black_and_white_tiled(img, opts, engine)
    heff = FM.film_optics(img, engine)
    tiles, crystals, developed, clumps = [], 0, 0, 0
    nx = plan["tiles_x"]
    for k, (xs, ys, tw, th) in enumerate(O.film_tile_spans(w, h, plan["tile_step"], plan["pad"])):
        tx, ty = k % nx, k // nx
        seed = (opts.seed + k * 73856093 + tx * 19349663 + ty * 83492791) % (1 << 64)
        p = F.GpuTileParams.for_preset(opts.preset, plan["grain_size_px"], tw, th, exposure=opts.exposure,
                                       tile_seed=seed)
        tiles.append(engine.render_tile_all_gpu(heff[ys:ys + th, xs:xs + tw], p))
        s = engine.last_tile_stats
        crystals += s["crystals"]
        developed += s["developed"]
        clumps += s["clumps"]
    tiles = np.concatenate(tiles)
    out2, _ = FM.film_composite(img, heff, tiles, opts, engine)
    assert (out == out2).all()
    assert rep.total_tiles == plan["tiles_x"] * plan["tiles_y"]
    assert (rep.total_crystals, rep.developed_crystals, rep.total_clumps) == (crystals, developed, clumps)
    # and the oracle composites the same tiles to within one level on a few pixels
    want, _ = O.film_composite(img, heff, tiles, tile_step=plan["tile_step"], pad=plan["pad"],
                               grain_size_px=plan["grain_size_px"], fold=1)
    d = np.abs(out.astype(np.int16) - want.astype(np.int16))
    assert d.max() <= 1 and (d > 0).mean() <= 1e-3
    # plain frame mode renders the same print; so does an RGB input of equal luma
    _, out3 = FM.render_film_black_and_white_tiled(img, FM.FilmRenderOptions(tile_size=64, seed=7,
                                                                              frame_mode="plain"), engine)
    assert (out3 == out).all()
    rgb = np.repeat(img[:, :, None], 3, axis=2)
    gray = O.film_luma(rgb, "tri-x")
    _, out4 = FM.render_film_black_and_white_tiled(rgb, opts, engine)
    _, out5 = FM.render_film_black_and_white_tiled(gray, opts, engine)
    assert (out4 == out5).all()


@pytest.mark.skip(reason=COMPOSITE_OFF)
@pytest.mark.parametrize("size", [2.4, 10.0, 20.0])
def test_letterbox_and_shadow_zero_variance(engine, size):
    """image_pipeline.rs:1235-1277 through the whole GPU render."""
    img = np.full((64, 64), 200, np.uint8)
    img[:8] = 0
    img[56:] = 0
    img[24:40, 24:40] = 0
    _, out = FM.render_film_black_and_white_tiled(
        img, FM.FilmRenderOptions(tile_size=64, grain_size_px=size, scanner_otf=True), engine)
    assert (out[:8] == 0).all() and (out[56:] == 0).all() and (out[26:38, 26:38] == 0).all()


@pytest.mark.skip(reason=COMPOSITE_OFF)
def test_highlight_fusion_burnout(engine):
    """image_pipeline.rs:1279-1326 through the whole GPU render."""
    img = np.full((64, 64), 128, np.uint8)
    img[:, :32] = 255
    _, out = FM.render_film_black_and_white_tiled(img, FM.FilmRenderOptions(tile_size=64, grain_size_px=2.4),
                                                  engine)
    fog = out[4:60, 4:28]
    assert fog.min() >= 253 and fog.max() == 255


@pytest.mark.skip(reason=COMPOSITE_OFF)
def test_film_render_is_deterministic_and_device_synthesis_runs(engine):
    img = _scene(120, 90, 3)
    opts = FM.FilmRenderOptions(tile_size=64, grain_size_px=1.5)
    r1, a = FM.render_film_black_and_white_tiled(img, opts, engine)
    r2, b = FM.render_film_black_and_white_tiled(img, opts, engine)
    assert (a == b).all() and r1.total_crystals == r2.total_crystals
    assert r1.total_crystals > 0 and r1.developed_crystals > 0
    rd, c = FM.render_film_black_and_white_tiled(img, FM.FilmRenderOptions(tile_size=64, grain_size_px=1.5,
                                                                         synthesis="device"), engine)
    assert rd.total_crystals > 0 and c.shape == a.shape
    # different crystals, same tone: the mean print level agrees
    assert abs(float(c.mean()) - float(a.mean())) < 2.0


@pytest.mark.skip(reason=COMPOSITE_OFF)
def test_overlapped_frame_blend_matches_oracle(engine):
    """fg_frame_read of an overlapped frame is the Hann blend of fg_frame_read_tiles."""
    from fgrain_amd import workloads as W
    cfg = W.load_config("tile320_trix_cubic")
    w, h, tile, ov = 200, 150, 64, 12
    exp = W.ramp_exposure(w, h, 0.05, 1.0)
    fr = W.Frame(engine, dict(cfg, frame={"width": w, "height": h, "tile": tile},
                              scan=dict(cfg["scan"], otf=0, supersample=1)), exp, overlap=ov)
    fr.render(W.Frame.BATCHED)
    tiles = fr.read_tiles()
    spans = O.film_tile_spans(w, h, tile, ov)
    assert tiles.size == sum(tw * th for _, _, tw, th in spans)
    assert [r[:4] for r in fr.tile_ranges()] == [tuple(s) for s in spans]
    _, _, ob = O.film_composite(np.full((h, w), 128, np.uint8), np.zeros((h, w), np.float32), tiles,
                                tile_step=tile, pad=ov, grain_size_px=2.2, intermediates=True)
    assert (_bits(fr.read()) == _bits(ob["d_norm"])).all()
    fr.render(W.Frame.PLAIN)
    assert (_bits(fr.read_tiles()) == _bits(tiles)).all()


def test_switched_off_entry_points_raise(engine):
    img = np.full((64, 64), 128, np.uint8)
    with pytest.raises(F.GpuError, match="switched off"):
        FM.render_film_black_and_white_tiled(img, FM.FilmRenderOptions(tile_size=64), engine)
    opts = FM.FilmRenderOptions(tile_size=64)
    with pytest.raises(F.GpuError, match="switched off"):
        FM.film_composite(img, np.zeros((64, 64), np.float32), np.zeros(64 * 64, np.float32), opts, engine)


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_overlapped_frame_tiles_equal_single_tile_frames(engine, mode):
    """An overlapped frame (image_pipeline.rs:397-434 layout) renders every padded tile
    exactly as a one-tile frame of that tile's extent, exposure crop and seed."""
    from fgrain_amd import workloads as W
    cfg = dict(W.load_config("tile320_trix_cubic"))
    cfg["scan"] = dict(cfg["scan"], otf=0, supersample=1)
    w, h, tile, ov = 150, 110, 64, 12
    exp = W.ramp_exposure(w, h, 0.05, 1.0)
    m = W.Frame.BATCHED if mode == "batched" else W.Frame.PLAIN
    fr = W.Frame(engine, dict(cfg, frame={"width": w, "height": h, "tile": tile}), exp, overlap=ov)
    fr.render(m)
    tiles = fr.read_tiles()
    spans = O.film_tile_spans(w, h, tile, ov)
    assert [r[:4] for r in fr.tile_ranges()] == [tuple(sp) for sp in spans]
    assert tiles.size == sum(tw * th for _, _, tw, th in spans)
    with pytest.raises(F.GpuError, match="switched off"):
        fr.read()
    off = 0
    for (xs, ys, tw, th, seed) in fr.tile_ranges():
        one = W.Frame(engine, dict(cfg, seed=seed, frame={"width": tw, "height": th, "tile": max(tw, th)}),
                      np.ascontiguousarray(exp[ys:ys + th, xs:xs + tw]))
        one.render(m)
        want = one.read().reshape(-1)
        assert (_bits(tiles[off:off + tw * th]) == _bits(want)).all(), (xs, ys)
        off += tw * th
