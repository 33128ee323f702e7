"""The differentiation workloads at their stated sizes on the GPU.

C4 (tile512_trix_adjoint: 512 x 512 px Tri-X tile, 256 packets per pixel, 67 M packets):
the path-replay gradients of mu_abs, R and the exposure map against central differences
with common random numbers, each to 1 % (the score-function gradient of g is checked on
a smaller tile over 400 seeds, test_gpu_adjoint.py: its noise per seed hardly falls with
the tile's area, so many small tiles reach 1 % at a twelfth of the cost).

C5 (fit1024_trix1600_dye3: 1024 x 1024 px, 3 dye layers of 16 tiles of 256 px, 128
packets per pixel, Tri-X pushed to 1600): Adam steps of the inverse fit reduce the loss,
and one of its tiles' expected forward and VJP equal the FP64 oracle's.

No reference counterpart (the reference has no gradients): the checks are the FP64 oracle
of the same expected model and finite differences.
"""
import numpy as np
import pytest

import fgrain_amd as F
from fgrain_amd import workloads as W
from fgrain_amd.adjoint import ExpectedFilm, mse_loss
from fgrain_amd.fit import DyeFit
from helpers import oracle_crystals
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4(engine):
    cfg = W.load_config("tile512_trix_adjoint")
    w, h, seed = cfg["frame"]["width"], cfg["frame"]["height"], cfg["seed"]
    em = F.VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, float(w), float(h)), seed)
    E = W.workload_exposure(cfg) * np.float32(W.packet_weight(cfg))
    m = W.expected_model(cfg, w, h, seed)
    tgt, _ = ExpectedFilm(em, m, engine).forward(E * np.float32(2.0 ** cfg["loss"]["target_shift_stops"]))
    return cfg, em, E, m, tgt


def _loss(engine, em, m, E, tgt):
    out, _ = ExpectedFilm(em, m, engine).forward(E)
    return mse_loss(out, tgt)[0]


def test_c4_pathwise_gradients_match_central_differences(engine, c4):
    """mu_abs (h = 1e-3 relative), R (h = 5e-2: the loss change at 1e-3 is below the float
    resolution of the densities) and a random direction of the exposure map (1e-3), with
    the same photon streams on both sides, each within 1 % of the VJP."""
    cfg, em, E, m, tgt = c4
    film = ExpectedFilm(em, m, engine)
    out, dose = film.forward(E)
    L, G = mse_loss(out, tgt)
    gE, pg = film.vjp(E, G, dose=dose)
    assert L > 0
    mu, R = m.exposure.crystal_absorption_coeff, m.optics.antihalation_reflectance
    for name, val, adj, rel in (("mu_abs", mu, pg.crystal_absorption_coeff, 1e-3),
                                ("R", R, pg.antihalation_reflectance, 5e-2)):
        hh = rel * val
        key = "crystal_absorption_coeff" if name == "mu_abs" else "antihalation_reflectance"
        fd = (_loss(engine, em, m.with_params(**{key: val + hh}), E, tgt) -
              _loss(engine, em, m.with_params(**{key: val - hh}), E, tgt)) / (2 * hh)
        print(f"C4 {name}: fd {fd:.6e} adjoint {adj:.6e}", flush=True)
        assert fd == pytest.approx(adj, rel=0.01)
    d = np.random.default_rng(5).standard_normal(E.shape)
    Ep = (E * (1 + 1e-3 * d)).astype(np.float32)
    Em = (E * (1 - 1e-3 * d)).astype(np.float32)
    fd = _loss(engine, em, m, Ep, tgt) - _loss(engine, em, m, Em, tgt)
    adj = float((gE.astype(np.float64) * (Ep.astype(np.float64) - Em)).sum())
    print(f"C4 exposure direction: fd {fd:.6e} adjoint {adj:.6e}", flush=True)
    assert fd == pytest.approx(adj, rel=0.01)


def _orc_cfg(m, E, w, h, cfg):
    o, x = m.optics, m.exposure
    return O.expected_cfg(E.shape, W=float(w), H=float(h), D=cfg["emulsion"]["depth"], mu_a=o.absorption_coeff,
                          mu_s=o.scattering_coeff, g=o.anisotropy_g, n_refr=o.refractive_index,
                          r_ahu=o.antihalation_reflectance, od_ahu=o.antihalation_density,
                          mu_abs=x.crystal_absorption_coeff, theta_max=x.illumination_angle_rad,
                          max_bounces=x.max_bounces, ppp=x.photons_per_pixel, seed=m.expose_seed,
                          fe=m.development.filament_expansion, sigma=m.scan.silver_opacity, w_px=w, h_px=h,
                          spp=m.scan.samples_per_pixel, ss=1, scan_seed=m.scan_seed,
                          pf=m.development.clumping_proximity_factor)


def test_c5_fit_at_its_stated_size(engine):
    """fit1024_trix1600_dye3 as stated: 1024 x 1024 px, 3 dye layers (seeds +0, +10007,
    +20011, image_pipeline.rs:1048-1061) of 16 tiles of 256 px, 128 packets per pixel,
    tri_x_1600 development (fe 1.32, pf 1.35).  Three Adam steps lower a finite loss, and
    one interior tile of the second layer has its expected forward and VJP (at the hidden
    exposure) equal to the FP64 oracle's: dose total within 0.2 %, per-crystal doses at the
    99th percentile within 5 %, transmittance within 2e-3 on average, parameter gradients
    within 1 % and the exposure gradient within 1 % (relative L2)."""
    cfg = W.load_config("fit1024_trix1600_dye3")
    fit = DyeFit(engine, cfg)
    assert len(fit.layers) == 3 and all(len(layer) == 16 for layer in fit.layers)
    assert fit.width == fit.height == 1024 and fit.tile == 256
    assert cfg["exposure"]["photons_per_pixel"] == 128
    r = fit.run(steps=3)
    print({k: v for k, v in r.items() if k != "loss"}, flush=True)
    loss = np.array(r["loss"])
    assert np.isfinite(loss).all() and len(loss) == 3
    assert loss[1] < loss[0] and loss[2] < loss[1]
    # one tile of layer 1 against the FP64 oracle
    x0, y0, tw, th, s, em, film = fit.layers[1][5]
    hidden = fit.hidden_field()[1]
    E = np.ascontiguousarray(hidden[y0:y0 + th, x0:x0 + tw] * np.float32(fit.weight))
    m = film.model
    out, dose, prob = film.forward(E, want_prob=True)
    cs = oracle_crystals(em)
    ocfg = _orc_cfg(m, E, tw, th, cfg)
    thr = m.development.threshold_photons
    ref = O.expected_forward(cs, E, ocfg, thr, int(m.scan.mode))
    assert dose.sum() == pytest.approx(ref["dose"].sum(), rel=2e-3)
    rel = np.abs(dose - ref["dose"]) / np.maximum(ref["dose"], 1.0)
    assert np.quantile(rel, 0.99) < 0.05
    d = np.abs(out - ref["out"])
    print(f"C5 tile: dose rel p99 {np.quantile(rel, 0.99):.2e}, out abs mean {d.mean():.2e} max {d.max():.2e}",
          flush=True)
    assert d.mean() < 2e-3 and d.max() < 0.05
    G = np.random.default_rng(6).standard_normal(out.shape).astype(np.float32) * np.float32(1e-3)
    gE, pg = film.vjp(E, G, dose=dose)
    rv = O.expected_vjp(cs, E, ocfg, thr, int(m.scan.mode), G.astype(np.float64), dose.astype(np.float64))
    print("C5 tile grads gpu", pg.array(), "fp64", rv["grads"], flush=True)
    assert np.allclose(pg.array(), rv["grads"], rtol=0.01)
    e = np.linalg.norm(gE - rv["grad_exposure"]) / np.linalg.norm(rv["grad_exposure"])
    assert e < 0.01
