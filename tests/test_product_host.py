"""Host-side product checks on the CPU (no GPU calls): the C ABI library loads and
exports every declared symbol, host synthesis and the BVH build equal the oracle's,
the threaded device layout visits nodes in the reference's DFS order, records and
boxes are laid out as the kernels read them, and errors follow the GpuError
convention."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import fgrain_amd as F
from fgrain_amd import _lib
from fgrain_amd import workloads as W
from helpers import oracle_crystals, trix_config
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    hdr = open(os.path.join(ROOT, "include", "fgrain_hip.h")).read()
    names = set(re.findall(r"\b(fg_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 25
    L = C.CDLL(_lib.LIB_PATH)
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(_lib.SIGNATURES) >= names
    assert L.fg_abi_version() == 1


def test_ctypes_structs_match_the_c_header(tmp_path):
    """Every ctypes mirror has the size the C compiler gives the header's struct."""
    pairs = {
        "fg_optics": _lib.Optics, "fg_emulsion_config": _lib.EmulsionConfigC,
        "fg_exposure_config": _lib.ExposureConfigC, "fg_expose_stats": _lib.ExposeStatsC,
        "fg_development_config": _lib.DevelopmentConfigC,
        "fg_development_report": _lib.DevelopmentReportC, "fg_scan_config": _lib.ScanConfigC,
        "fg_tuning": _lib.TuningC, "fg_tile_params": _lib.TileParamsC,
        "fg_tile_stats": _lib.TileStatsC, "fg_frame_desc": _lib.FrameDescC,
        "fg_frame_stats": _lib.FrameStatsC, "fg_expected_model": _lib.ExpectedModelC,
        "fg_param_grads": _lib.ParamGradsC, "fg_adjoint_stats": _lib.AdjointStatsC,
        "fg_film_options": _lib.FilmOptionsC, "fg_film_plan": _lib.FilmPlanC,
        "fg_film_report": _lib.FilmReportC, "fg_composite_stats": _lib.CompositeStatsC,
        "fg_device_tile": _lib.DeviceTileC,
    }
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "fgrain_hip.h"\nint main(void){\n' +
                   "".join(f'printf("%zu\\n", sizeof({k}));\n' for k in pairs) + "return 0;}\n")
    exe = tmp_path / "sz"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                            check=True).stdout.split()]
    for (k, cls), n in zip(pairs.items(), sizes):
        assert C.sizeof(cls) == n, k


@pytest.mark.parametrize("morph", [0, 1, 2])
def test_host_synthesis_equals_oracle(morph):
    cfg = trix_config(48, 40, morph=morph, aspect=8.0 if morph == 1 else 1.0)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 17 + morph)
    c = em.crystals()
    oc = O.synthesize(W=48, H=40, D=24, r_min=cfg.r_min, mean_radius=cfg.mean_crystal_radius,
                      size_sigma=cfg.size_variance, morph=morph, aspect=cfg.tabular_aspect_ratio,
                      base_sens=cfg.base_sensitivity, sens_sigma=cfg.sensitivity_sigma,
                      min_radius=cfg.min_radius, min_half_thickness=cfg.min_half_thickness,
                      seed=17 + morph)
    assert len(em) == oc.n > 100
    for k in ("pos", "dims", "quat", "sens", "morph"):
        assert np.array_equal(c[k], getattr(oc, k)), k


def test_host_bvh_equals_oracle_node_by_node():
    em = F.VirtualEmulsionLayer.synthesize(trix_config(80, 64), 3)
    b, ob = em.bvh(), O.bvh_build(oracle_crystals(em))
    for k in b:
        assert np.array_equal(b[k], ob[k]), k


def _threaded_walk_hits(nodes, recs_idx, cs, o, d):
    """Walk the threaded layout exactly as the kernels do (float box tests) and return
    the crystals hit, in visit order, using the oracle's float intersection."""
    f = nodes.view(np.float32)
    o32 = o.astype(np.float32)
    d32 = d.astype(np.float32)
    with np.errstate(divide="ignore"):
        inv = (np.float32(1) / d32).astype(np.float32)
    i, out = 0, []
    nn = len(nodes)
    while i < nn:
        lo, hi = f[i, 0:3], f[i, 3:6]
        with np.errstate(invalid="ignore"):
            t1 = (lo - o32) * inv
            t2 = (hi - o32) * inv
        te = np.fmax(np.fmax(np.fmin(t1[0], t2[0]), np.fmin(t1[1], t2[1])), np.fmin(t1[2], t2[2]))
        tx = np.fmin(np.fmin(np.fmax(t1[0], t2[0]), np.fmax(t1[1], t2[1])), np.fmax(t1[2], t2[2]))
        hit = tx >= te and tx > 0
        prim = int(nodes[i, 7])
        if hit and prim != 0xFFFFFFFF:
            for k in range(prim >> 4, (prim >> 4) + (prim & 15)):
                cid = int(recs_idx[k])
                if O.crystal_intersect(cs, cid, o32, d32, fp32=True) is not None:
                    out.append(cid)
        # a hit node of either kind continues at i + 1 (a leaf's skip link is i + 1)
        i = i + 1 if hit else int(nodes[i, 6])
    return out


def test_host_development_equals_oracle():
    """fg_develop without an engine runs the host restatement (development.rs:59-188):
    flags, clump ids, boosted expansions and the report equal the oracle's exactly."""
    import ctypes as C
    from fgrain_amd import _lib as L
    em = F.VirtualEmulsionLayer.synthesize(trix_config(96, 96), 4)
    counts = np.random.default_rng(0).poisson(3.0, len(em)).astype(np.uint32)
    em.set_state(counts=counts)
    cfg = F.DevelopmentConfig(4, 1.2, 1.25).c()
    rep = L.DevelopmentReportC()
    L.check(L.lib().fg_develop(None, em._h, C.byref(cfg), C.byref(rep)))
    st = em.state()
    dev, fe, cid, orep = O.develop(oracle_crystals(em), counts, 4, 1.2, 1.25)
    assert (st["developed"] == dev).all() and (st["clump_id"] == cid).all() and (st["fe"] == fe).all()
    assert rep.developed_crystals == orep["developed"] and rep.total_clumps == orep["clumps"]
    assert rep.max_clump_size == orep["max_clump"]


def test_threaded_layout_reproduces_dfs_hit_order():
    em = F.VirtualEmulsionLayer.synthesize(trix_config(40, 40), 9)
    nodes, recs = em.device_layout()
    idx = em.bvh()["indices"]
    cs = oracle_crystals(em)
    rng = np.random.default_rng(4)
    o = np.stack([rng.uniform(0, 40, 60), rng.uniform(0, 40, 60), rng.uniform(0, 24, 60)], 1)
    d = rng.normal(size=(60, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nh, oidx, _, _ = O.all_hits(cs, o.astype(np.float32), d.astype(np.float32), fp32=True)
    total = 0
    for r in range(60):
        got = _threaded_walk_hits(nodes, idx, cs, o[r], d[r])
        assert got == oidx[r, :nh[r]].tolist()
        total += len(got)
    assert total > 100


def _down(x):
    f = np.float32(x)
    return f if float(f) <= x else np.nextafter(f, np.float32(-np.inf))


def _up(x):
    f = np.float32(x)
    return f if float(f) >= x else np.nextafter(f, np.float32(np.inf))


def test_device_records_and_outward_rounded_boxes():
    em = F.VirtualEmulsionLayer.synthesize(trix_config(32, 32), 2)
    nodes, recs = em.device_layout()
    b = em.bvh()
    c = em.crystals()
    rf = recs.view(np.float32)
    k = b["indices"]
    assert np.array_equal(rf[:, 0:3], c["pos"][k].astype(np.float32))
    assert np.array_equal(rf[:, 3], c["sens"][k].astype(np.float32))
    assert np.array_equal(rf[:, 4:8], c["quat"][k].astype(np.float32))
    assert np.array_equal(rf[:, 8:11], c["dims"][k].astype(np.float32))
    assert np.array_equal(recs[:, 11], c["morph"][k].astype(np.uint32))
    # re-derive the threaded slot of every reference node: right subtree directly after
    # its parent, left subtree after the right one; skip = next slot after the subtree
    nn = len(b["left"])
    size = np.ones(nn, int)
    for i in range(nn - 1, -1, -1):
        if b["count"][i] == 0:
            size[i] = 1 + size[b["left"][i]] + size[b["right"][i]]
    stack = [(0, 0, nn)]
    f = nodes.view(np.float32)
    seen = 0
    while stack:
        r, slot, skip = stack.pop()
        seen += 1
        for a in range(3):
            assert f[slot, a] == _down(b["bmin"][r, a])
            assert f[slot, 3 + a] == _up(b["bmax"][r, a])
        assert nodes[slot, 6] == skip
        if b["count"][r] > 0:
            assert nodes[slot, 7] == (b["offset"][r] << 4) | b["count"][r]
            assert skip == slot + 1   # the kernels' walk relies on this
        else:
            assert nodes[slot, 7] == 0xFFFFFFFF
            rs = slot + 1
            ls = rs + size[b["right"][r]]
            stack.append((b["right"][r], rs, ls))
            stack.append((b["left"][r], ls, skip))
    assert seen == nn == len(nodes)


def test_errors_follow_gpuerror_convention():
    cfg = trix_config(10, 10)
    cfg.r_min = 0.0
    with pytest.raises(F.GpuError) as e:
        F.VirtualEmulsionLayer.synthesize(cfg, 1)
    assert e.value.status == 7 and "r_min" in str(e.value)
    with pytest.raises(F.GpuError) as e:
        F.VirtualEmulsionLayer.from_crystals(trix_config(4, 4), [[1, 1, 1]], [[1, 1, 1]],
                                            [[1, 0, 0, 0]], 5, 1.0)
    assert e.value.status == 7


def test_no_device_reports_adapter_not_found():
    """GpuEngine::new without an adapter -> AdapterNotFound; get_gpu_engine caches the
    failure and returns None (gpu.rs:9-14, image_pipeline.rs:391-395)."""
    if F.get_gpu_engine() is not None:
        pytest.skip("a HIP device is present")
    with pytest.raises(F.GpuError) as e:
        F.GpuEngine.new()
    assert e.value.status == 1
    assert F.get_gpu_engine() is None


def test_tile_params_follow_the_reference_tables():
    """image_pipeline.rs:455-515 (tile geometry, spp, threshold) and 924-1002 (the CPU
    tile path's emulsion, exposure, development and scan values)."""
    p = F.GpuTileParams.for_preset("tri-x", 2.4, 384, 384, exposure=1.0, push=0.0, tile_seed=7).c
    assert p.film_w == pytest.approx(384 / 2.4) and p.depth == 14.0 and p.r_min == 1.2
    assert p.samples_per_pixel == 4 and p.threshold_photons == 4 and p.max_bounces == 48
    assert p.photons_per_sq_unit == 6.0 and p.illumination_angle_rad == 0.12
    assert p.crystal_absorption_coeff == 1.3 and p.silver_opacity == 0.25
    assert p.filament_expansion == pytest.approx(1.2) and p.clumping_proximity_factor == pytest.approx(1.25)
    p = F.GpuTileParams.for_preset("tri-x", 2.4, 10, 10, exposure=0.01, push=2.0).c
    assert p.depth == 14.5 and p.base_sensitivity == 1.35 and p.photons_per_sq_unit == pytest.approx(0.3)
    assert p.filament_expansion == pytest.approx(1.32)
    p = F.GpuTileParams.for_preset("t-max", 2.4, 10, 10).c
    assert p.morphology == 1 and p.tabular_aspect_ratio == 6.5 and p.optics.antihalation_density == 1.10
    p = F.GpuTileParams.for_preset("hp5", 2.4, 10, 10, push=1.0).c
    assert p.morphology == 2 and p.filament_expansion == pytest.approx(1.20)


def test_workload_exposures():
    cfg = W.load_config("frame24mp_trix_cubic")
    e = W.blob_exposure(600, 400, 0)
    assert e.shape == (400, 600) and e.dtype == np.float32
    assert e.min() >= 0.01 - 1e-7 and e.max() <= 4.0 + 1e-6
    lv = np.round((e - 0.01) / 3.99 * 255)
    assert np.allclose((e - 0.01) / 3.99 * 255, lv, atol=1e-3)  # 8-bit levels
    r = W.ramp_exposure(320, 2, 0.01, 4.0)
    assert r[0, 0] == np.float32(0.01) and r[1, -1] == np.float32(4.0)
    assert cfg["frame"]["width"] == 6000 and cfg["exposure"]["photons_per_pixel"] == 128


def test_fit_tile_seeds_follow_the_tiled_renderer():
    """The fit's per-tile seeds are the frame path's (image_pipeline.rs:430-434)."""
    from fgrain_amd.fit import tile_seed
    cfg = W.load_config("frame24mp_trix_cubic")
    fr = W.Frame.__new__(W.Frame)
    fr.seed, fr.tile, fr.width, fr.height = cfg["seed"], 320, 1000, 700
    seeds = [t[4] for t in W.Frame.tile_ranges(fr)]
    idx = 0
    ours = []
    for ty in range(3):
        for tx in range(4):
            ours.append(tile_seed(cfg["seed"], idx, tx, ty))
            idx += 1
    assert seeds == ours


def test_smooth_field_range_and_packet_weight():
    f = W.smooth_field(64, 48, 3, 0.05, 3.0, 6)
    assert f.shape == (48, 64) and f.min() == pytest.approx(0.05) and f.max() == pytest.approx(3.0)
    assert np.array_equal(f, W.smooth_field(64, 48, 3, 0.05, 3.0, 6))
    cfg = W.load_config("tile512_trix_adjoint")
    assert W.packet_weight(cfg) == pytest.approx(0.10417 / 256)
    assert W.packet_weight(W.load_config("tile320_trix_cubic")) == 1.0


@pytest.mark.parametrize("kw", [
    dict(), dict(grain_size_px=0.5), dict(grain_size_px=0.05, tile_size=1000), dict(grain_size_px=15.0),
    dict(grain_size_px=2.2, auto_scale=True, reference_res=3000), dict(scale=2.0, grain_strength=0.7),
    dict(preset="tri-x-1600", push=0.5), dict(preset="t-max", push=1.0), dict(preset="hp5"),
])
def test_film_plan_follows_the_reference(kw):
    """fg_film_plan_for (host logic) against the oracle's restatement of
    image_pipeline.rs:346-385 and :651-658."""
    from fgrain_amd import film as FM
    opts = FM.FilmRenderOptions(**kw)
    for w, h in [(6000, 4000), (640, 480), (97, 61)]:
        p = FM.film_plan(opts, w, h)
        o = O.film_plan(w, h, grain_size_px=opts.grain_size_px, grain_strength=opts.grain_strength,
                        scale=opts.scale, auto_scale=opts.auto_scale, reference_res=opts.reference_res,
                        tile_size=opts.tile_size)
        assert p["grain_size_px"] == o["grain_size_px"] and p["strength"] == o["strength"]
        assert (p["tile_step"], p["pad"], p["tiles_x"], p["tiles_y"]) == \
            (o["tile_step"], o["pad"], o["tiles_x"], o["tiles_y"])
        assert p["push"] == O.film_effective_push(opts.preset, opts.push)
        assert p["base_nominal"] == pytest.approx(O.film_base_nominal(opts.preset))
        # the Gaussian of the low-frequency separation, as the oracle composites it
        _, st = O.film_composite(np.zeros((8, 8), np.uint8), np.zeros((8, 8), np.float32),
                                 np.zeros(64, np.float32), tile_step=48, pad=12,
                                 grain_size_px=p["grain_size_px"])
        assert p["filter_sigma"] == st["filter_sigma"] and p["k_rad"] == st["k_rad"]


def test_film_options_validation():
    from fgrain_amd import film as FM
    with pytest.raises(F.GpuError):
        FM.film_plan(FM.FilmRenderOptions(tile_size=40), 100, 100)
    with pytest.raises(F.GpuError):
        FM.film_plan(FM.FilmRenderOptions(grain_size_px=0.0), 100, 100)
    with pytest.raises(F.GpuError):
        FM.film_plan(FM.FilmRenderOptions(), 0, 100)
    with pytest.raises(NotImplementedError):
        FM.render_film_black_and_white_tiled(np.zeros((8, 8), np.uint8), FM.FilmRenderOptions(color=True))
    with pytest.raises(NotImplementedError):
        FM.render_film_black_and_white_tiled(np.zeros((8, 8), np.uint8), FM.FilmRenderOptions(supersample=2))
