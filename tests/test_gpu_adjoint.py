"""The expected model and its adjoint on the GPU (fg_expected_forward, fg_vjp).

No reference counterpart: the GPU path is checked against the FP64 oracle of the same
model (same Philox streams, f64 arithmetic), against central finite differences of the
GPU forward with common random numbers (identical keys), and through the VJP interface
(ExpectedFilm and the torch.autograd.Function adapter).

Exposure maps are given as packet weights: relative exposure times the photon density of
the reference tile path (6 photons/um^2 at exposure 1, 1.0417 per px^2 at 2.4 px/um,
image_pipeline.rs:942-947) over the packets per pixel, so doses straddle the threshold.
"""
import numpy as np
import pytest

import fgrain_amd as F
from fgrain_amd.adjoint import ExpectedFilm, ExpectedFilmFunction, ExpectedModel, mse_loss
from oracle import oracle as O

from helpers import oracle_crystals, trix_config

pytestmark = pytest.mark.gpu

PHOTONS_PER_PX2 = 1.0417


def _field(h, w, ppp, seed=0):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    f = np.zeros((h, w))
    for _ in range(6):
        cx, cy, s = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(0.15, 0.4) * max(w, h)
        f += rng.uniform(0.3, 1.0) * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    f = (f - f.min()) / (f.max() - f.min())
    return ((0.05 + 2.95 * f) * PHOTONS_PER_PX2 / ppp).astype(np.float32)


def _model(w, h, ppp, seed=11, mu_abs=0.9, g=0.85, R=0.14, mode=F.ScanMode.NegativeDensity):
    optics = F.GelatinOptics(1.534, 0.00125, 0.0208333333, g, R, 0.8)
    return ExpectedModel(F.ExposureConfig(0.0, 0.12, mu_abs, 48, ppp, 0.0), optics,
                         F.DevelopmentConfig(4, 1.2, 1.25), F.ScanConfig(w, h, 4, mode, 0.1041666667, 1, False),
                         expose_seed=seed, scan_seed=12)


def _orc_cfg(E, m: ExpectedModel, w, h):
    o = m.optics
    return O.expected_cfg(E.shape, W=float(w), H=float(h), D=24.0, mu_a=o.absorption_coeff,
                          mu_s=o.scattering_coeff, g=o.anisotropy_g, n_refr=o.refractive_index,
                          r_ahu=o.antihalation_reflectance, od_ahu=o.antihalation_density,
                          mu_abs=m.exposure.crystal_absorption_coeff, theta_max=0.12, max_bounces=48,
                          ppp=m.exposure.photons_per_pixel, seed=m.expose_seed, fe=1.2,
                          sigma=0.1041666667, w_px=w, h_px=h, spp=4, ss=1, scan_seed=m.scan_seed,
                          pf=m.development.clumping_proximity_factor)


@pytest.fixture(scope="module")
def tile():
    w = h = 64
    cfg = trix_config(w, h, r_ahu=0.14)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 7)
    ppp = 32
    E = _field(h, w, ppp)
    return em, E, w, h, ppp


def _target(engine, em, E, w, h, ppp):
    out, _ = ExpectedFilm(em, _model(w, h, ppp, seed=99), engine).forward(E * np.float32(2 ** 0.3))
    return out


def test_expected_forward_matches_fp64_oracle(engine, tile):
    em, E, w, h, ppp = tile
    m = _model(w, h, ppp)
    out, dose, prob = ExpectedFilm(em, m, engine).forward(E, want_prob=True)
    cs = oracle_crystals(em)
    ref = O.expected_forward(cs, E, _orc_cfg(E, m, w, h), 4, 3)
    # f32 paths drift from the f64 paths on a few packets; totals and per-crystal values
    # agree to far below the Monte Carlo noise of the estimate
    assert dose.sum() == pytest.approx(ref["dose"].sum(), rel=2e-3)
    rel = np.abs(dose - ref["dose"]) / np.maximum(ref["dose"], 1.0)
    print("dose rel err: median %.2e p99 %.2e max %.2e" % (np.median(rel), np.quantile(rel, 0.99), rel.max()))
    assert np.quantile(rel, 0.99) < 0.05
    assert np.abs(prob - ref["prob"]).mean() < 2e-3
    d = np.abs(out - ref["out"])
    print("out abs err: mean %.2e max %.2e" % (d.mean(), d.max()))
    assert d.mean() < 2e-3 and d.max() < 0.05
    assert 0.05 < prob.mean() < 0.95          # the workload straddles the threshold


def test_vjp_matches_fp64_oracle_gradient(engine, tile):
    em, E, w, h, ppp = tile
    m = _model(w, h, ppp)
    film = ExpectedFilm(em, m, engine)
    out, dose = film.forward(E)
    tgt = _target(engine, em, E, w, h, ppp)
    L, G = mse_loss(out, tgt)
    gE, pg, gd = film.vjp(E, G, dose=dose, want_grad_dose=True)
    cs = oracle_crystals(em)
    ref = O.expected_vjp(cs, E, _orc_cfg(E, m, w, h), 4, 3, G.astype(np.float64), dose.astype(np.float64))
    print("grads gpu", pg.array(), "fp64", ref["grads"])
    assert np.allclose(pg.array(), ref["grads"], rtol=0.01)
    e = np.linalg.norm(gE - ref["grad_exposure"]) / np.linalg.norm(ref["grad_exposure"])
    print("grad_exposure rel L2 %.2e" % e)
    assert e < 0.01
    eg = np.linalg.norm(gd - ref["grad_dose"]) / np.linalg.norm(ref["grad_dose"])
    assert eg < 1e-3


def _loss(engine, em, m, E, tgt):
    out, _ = ExpectedFilm(em, m, engine).forward(E)
    return mse_loss(out, tgt)[0]


def test_pathwise_gradients_match_central_differences(engine, tile):
    """mu_abs, R and a random exposure direction with common random numbers: h = 1e-3
    relative for mu_abs and E; for R, whose loss change at that step is below the float
    resolution of the scanned densities, h = 5e-2 relative (the dose is a polynomial in R
    whose quadratic term, paths reflected twice, is R_eff^2 ~ 1e-5 of the linear one)."""
    em, E, w, h, ppp = tile
    m = _model(w, h, ppp)
    tgt = _target(engine, em, E, w, h, ppp)
    film = ExpectedFilm(em, m, engine)
    out, dose = film.forward(E)
    gE, pg = film.vjp(E, mse_loss(out, tgt)[1], dose=dose)
    for name, val, adj in (("mu_abs", 0.9, pg.crystal_absorption_coeff), ("R", 0.14, pg.antihalation_reflectance)):
        hh = (1e-3 if name == "mu_abs" else 5e-2) * val
        kw_p = {"crystal_absorption_coeff" if name == "mu_abs" else "antihalation_reflectance": val + hh}
        kw_m = {"crystal_absorption_coeff" if name == "mu_abs" else "antihalation_reflectance": val - hh}
        fd = (_loss(engine, em, m.with_params(**kw_p), E, tgt) - _loss(engine, em, m.with_params(**kw_m), E, tgt)) / (2 * hh)
        print(name, "fd %.6e adjoint %.6e" % (fd, adj))
        assert fd == pytest.approx(adj, rel=0.01)
    d = np.random.default_rng(3).standard_normal(E.shape)
    Ep = (E * (1 + 1e-3 * d)).astype(np.float32)
    Em = (E * (1 - 1e-3 * d)).astype(np.float32)
    fd = _loss(engine, em, m, Ep, tgt) - _loss(engine, em, m, Em, tgt)
    adj = float((gE.astype(np.float64) * (Ep.astype(np.float64) - Em)).sum())
    print("E direction fd %.6e adjoint %.6e" % (fd, adj))
    assert fd == pytest.approx(adj, rel=0.01)


def test_hg_score_gradient_agrees_with_finite_differences(engine):
    """Score-function dL/dg against central differences with common random numbers, to 1 %.

    The difference quotient has a visible O(h^2) truncation term (it moves by about 2 %
    between h = 3 % and 10 % of g), so it is extrapolated to h -> 0 from both steps on
    the same photon streams (Richardson: FD0 = FD(h1) - (FD(h2) - FD(h1)) / ((h2/h1)^2 - 1)).
    400 photon-stream seeds on a 128 x 128 tile at 256 packets per pixel make three
    standard errors of score - FD0 less than 1 % of FD0."""
    w = h = 128
    ppp = 256
    em = F.VirtualEmulsionLayer.synthesize(trix_config(w, h, r_ahu=0.14), 8)
    E = _field(h, w, ppp, seed=1)
    tgt = _target(engine, em, E, w, h, ppp)
    g, h1, h2 = 0.85, 0.03, 0.10
    sc, fd0, fd1 = [], [], []
    for seed in range(400):
        m = _model(w, h, ppp, seed=1000 + seed)
        film = ExpectedFilm(em, m, engine)
        out, dose = film.forward(E)
        sc.append(film.vjp(E, mse_loss(out, tgt)[1], dose=dose)[1].anisotropy_g)
        f = [(_loss(engine, em, m.with_params(anisotropy_g=g * (1 + r)), E, tgt) -
              _loss(engine, em, m.with_params(anisotropy_g=g * (1 - r)), E, tgt)) / (2 * g * r) for r in (h1, h2)]
        fd1.append(f[0])
        fd0.append(f[0] - (f[1] - f[0]) / ((h2 / h1) ** 2 - 1))
    sc, fd0, fd1 = np.array(sc), np.array(fd0), np.array(fd1)
    dev = sc - fd0   # paired: the same streams on both sides
    se = dev.std(ddof=1) / np.sqrt(len(dev))
    ref = abs(fd0.mean())
    print("g score %.5e  fd(h=3%%) %.5e  fd(h->0) %.5e  score - fd0 = %.3f%% +- %.3f%% (1 se)"
          % (sc.mean(), fd1.mean(), fd0.mean(), 100 * dev.mean() / ref, 100 * se / ref), flush=True)
    assert 3 * se < 0.01 * ref
    assert abs(dev.mean()) < 0.01 * ref


def test_vjp_recomputes_dose_and_is_reproducible(engine, tile):
    em, E, w, h, ppp = tile
    m = _model(w, h, ppp)
    film = ExpectedFilm(em, m, engine)
    out, dose = film.forward(E)
    G = np.random.default_rng(4).standard_normal(out.shape).astype(np.float32)
    gE1, p1 = film.vjp(E, G, dose=dose)
    gE2, p2 = film.vjp(E, G)
    gE3, p3 = film.vjp(E, G, dose=dose)
    assert np.allclose(p1.array(), p3.array(), rtol=1e-12)
    assert np.allclose(gE1, gE3, rtol=1e-6, atol=0)
    assert np.allclose(p1.array(), p2.array(), rtol=1e-6)
    assert np.allclose(gE1, gE2, rtol=1e-5, atol=1e-30)
    out2, dose2 = film.forward(E)
    assert np.allclose(dose, dose2, rtol=1e-6) and np.allclose(out, out2, atol=1e-6)


def test_torch_autograd_function_matches_vjp(engine, tile):
    import torch
    em, E, w, h, ppp = tile
    m = _model(w, h, ppp)
    film = ExpectedFilm(em, m, engine)
    tgt = torch.from_numpy(_target(engine, em, E, w, h, ppp))
    Et = torch.from_numpy(E.copy()).requires_grad_(True)
    params = torch.tensor(m.params(), dtype=torch.float64, requires_grad=True)
    out = ExpectedFilmFunction.apply(Et, params, film)
    loss = ((out.double() - tgt.double()) ** 2).mean()
    loss.backward()
    o2, dose = film.forward(E)
    gE, pg = film.vjp(E, mse_loss(o2, tgt.numpy())[1], dose=dose)
    assert np.allclose(Et.grad.numpy(), gE, rtol=1e-5, atol=1e-30)
    assert np.allclose(params.grad.numpy(), pg.array(), rtol=1e-5)


def test_model_validation_errors(engine, tile):
    em, E, w, h, ppp = tile
    m = _model(w, h, ppp)
    m.scan.otf = True
    with pytest.raises(F.GpuError, match="otf"):
        ExpectedFilm(em, m, engine).forward(E)
    m = _model(w, h, 0)
    with pytest.raises(F.GpuError, match="photons_per_pixel"):
        ExpectedFilm(em, m, engine).forward(E)
    m = _model(w, h, ppp, g=1.0)
    with pytest.raises(F.GpuError, match="anisotropy"):
        ExpectedFilm(em, m, engine).forward(E)
    with pytest.raises(ValueError):
        ExpectedFilm(em, _model(w, h, ppp), engine).vjp(E, np.zeros((3, 3), np.float32))


def test_persistent_replay_equals_thread_per_packet(engine, tile):
    """Both replay kernels follow identical packet paths; only the order of the double
    accumulations differs."""
    em, E, w, h, ppp = tile
    m = _model(w, h, ppp)
    G = np.random.default_rng(5).standard_normal((h, w)).astype(np.float32)
    res = []
    old = engine.tuning()
    try:
        for kernel in (0, 1):
            engine.set_tuning(transport_kernel=kernel)
            film = ExpectedFilm(em, m, engine)
            out, dose = film.forward(E)
            st = dict(film.last_stats)
            gE, pg, gd = film.vjp(E, G, dose=dose, want_grad_dose=True)
            res.append((out, dose, gE, pg.array(), gd, st))
    finally:
        engine.set_tuning(**old)
    a, b = res
    assert np.allclose(a[1], b[1], rtol=1e-6, atol=1e-6)
    assert np.abs(a[0] - b[0]).max() <= 1e-5
    assert np.allclose(a[2], b[2], rtol=1e-5, atol=1e-12)
    assert np.allclose(a[3], b[3], rtol=1e-5)
    assert a[5]["bounces"] == b[5]["bounces"] and a[5]["events"] == b[5]["events"]
