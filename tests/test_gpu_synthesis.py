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
