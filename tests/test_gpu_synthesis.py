"""Device emulsion synthesis (fg_emulsion_synthesize_device, frames with synthesis
"device") against its oracle twin: the same points and crystals, and a frame built from
device-synthesised tiles renders identically on the plain and batched paths."""
import copy
import time

import numpy as np
import pytest

import fgrain_amd as F
from fgrain_amd import workloads as W
from helpers import trix_config
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("morph", [0, 1, 2])
def test_device_synthesis_equals_oracle_twin(engine, morph):
    cfg = trix_config(96, 80, morph=morph, aspect=8.0 if morph == 1 else 1.0)
    em = F.VirtualEmulsionLayer.synthesize_device(cfg, 17, engine)
    c = em.crystals()
    W_, H_, D_ = cfg.dimensions
    ref = O.synthesize_darts(W=W_, H=H_, D=D_, r_min=cfg.r_min, mean_radius=cfg.mean_crystal_radius,
                             size_sigma=cfg.size_variance, morph=morph, aspect=cfg.tabular_aspect_ratio,
                             base_sens=cfg.base_sensitivity, sens_sigma=cfg.sensitivity_sigma,
                             min_radius=cfg.min_radius, min_half_thickness=cfg.min_half_thickness,
                             k_samples=cfg.k_samples, seed=17)
    assert len(em) == ref.n > 500
    assert np.array_equal(c["pos"], ref.pos)
    assert np.array_equal(c["dims"], ref.dims)
    assert np.array_equal(c["quat"], ref.quat)
    assert np.array_equal(c["sens"], ref.sens)
    # the BVH is the reference's build over these crystals
    bvh = em.bvh()
    ob = O.bvh_build(O.Crystals(c["pos"], c["dims"], c["quat"], c["morph"], c["sens"]))
    assert np.array_equal(bvh["indices"], ob["indices"])


def test_device_synthesised_frame_renders_identically_on_both_paths(engine):
    cfg = copy.deepcopy(W.load_config("frame24mp_trix_cubic"))
    cfg["exposure"]["photons_per_pixel"] = 8
    w, h = 700, 420
    exp = W.blob_exposure(w, h, 5, sigma=(40, 160))
    t = time.perf_counter()
    fd = W.Frame(engine, cfg, exp, frame=(w, h), synthesis="device")
    td = time.perf_counter() - t
    t = time.perf_counter()
    fh = W.Frame(engine, cfg, exp, frame=(w, h), synthesis="host")
    th = time.perf_counter() - t
    print(f"frame synthesis: device {td:.2f} s ({fd.crystals} crystals), host {th:.2f} s ({fh.crystals})")
    assert fd.crystals > fh.crystals          # saturated dart throwing is denser than Bridson
    sp = fd.render(W.Frame.PLAIN)
    Dp, cp = fd.read(), fd.counts()
    sb = fd.render(W.Frame.BATCHED)
    assert np.array_equal(cp, fd.counts())
    assert np.abs(Dp - fd.read()).max() <= 1e-5
    assert sp["developed"] == sb["developed"] > 0


@pytest.mark.parametrize("workload", ["frame24mp_trix_cubic", "frame24mp_tmax_tabular"])
def test_device_built_frame_tiles_equal_host_builds(engine, workload):
    """Frames synthesised on the device get every tile's crystals, transport BVH and
    records from the device (compaction, attributes, BVH::build of bvh.rs:157-295 with the
    reference's partition order): per tile, the threaded nodes and leaf-order records equal
    word for word the host build over the same crystals (build_ref_bvh and its layout;
    T-Max tiles with uneven densities and tabular crystals), and the crystals are the oracle
    twin's darts with the attribute draws of emulsion.rs:160-207 (f64 transcendentals of
    the device's libm: within 1e-13 relative of the host's)."""
    cfg = copy.deepcopy(W.load_config(workload))
    cfg["exposure"]["photons_per_pixel"] = 4
    w, h = 700, 420   # 3 x 2 tiles, partial edge tiles
    fr = W.Frame(engine, cfg, W.blob_exposure(w, h, 3, sigma=(40, 160)), frame=(w, h), synthesis="device")
    mult = fr._mult
    e = cfg["emulsion"]
    for k, (x0, y0, tw, th, seed) in enumerate(fr.tile_ranges()):
        t = fr.tile_data(k)
        assert len(t["pos"]) > 1000
        ec = W.emulsion_config(cfg, float(tw), float(th))
        em = F.VirtualEmulsionLayer.from_crystals(ec, t["pos"], t["dims"], t["quat"], t["morph"], t["sens"])
        nodes, recs = em.device_layout()
        assert nodes.shape == t["nodes"].shape and np.array_equal(nodes, t["nodes"]), k
        assert np.array_equal(recs, t["records"]), k
        r_min = e["r_min"] / np.cbrt(max(float(mult[k]), 1e-3)) if mult is not None else e["r_min"]
        ref = O.synthesize_darts(W=float(tw), H=float(th), D=e["depth"], r_min=r_min,
                                 mean_radius=e["mean_crystal_radius"], size_sigma=e["size_variance"],
                                 morph=e["morphology"], aspect=e["tabular_aspect_ratio"],
                                 base_sens=e["base_sensitivity"], sens_sigma=e["sensitivity_sigma"],
                                 min_radius=e["min_radius"], min_half_thickness=e["min_half_thickness"],
                                 k_samples=e["k_samples"], seed=seed)
        assert np.array_equal(t["pos"], ref.pos), k
        for key, v in (("dims", ref.dims), ("quat", ref.quat), ("sens", ref.sens)):
            np.testing.assert_allclose(t[key], v, rtol=1e-13, atol=1e-15, err_msg=f"{key} tile {k}")
    st = fr.render(W.Frame.BATCHED)
    assert st["graph_used"] == 1 and 0 < st["developed"] < st["crystals"]
