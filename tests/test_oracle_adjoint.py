"""Pins of the oracle's expected (differentiable) model -- run on the CPU.

The expected model has no reference counterpart, so it is pinned by mathematics:
scipy's Poisson tail, quadrature identities of the Henyey-Greenstein score, closed forms
for one crystal, Euler's identity for the dose (homogeneous of degree 1 in the exposure
map), and central finite differences of the FP64 forward with common random numbers
(identical Philox keys).  A plausible slip (a dropped tangent term, a wrong sign in the
score, a missing 1 - pa factor, a mis-scaled ray weight) fails at least one of them.
"""
import numpy as np
import pytest
from scipy import integrate
from scipy.stats import poisson

from oracle import oracle as O

TRIX_OPTICS = dict(mu_a=0.00125, mu_s=0.0208333333, n_refr=1.534, od_ahu=0.8)


def _one(pos, a, sens=1.0):
    return O.Crystals(np.array([pos], float), np.array([[a, a, a]], float),
                      np.array([[1.0, 0, 0, 0]]), np.zeros(1, np.int32), np.array([sens]))


def _tile(w=40, seed=3):
    return O.synthesize(W=w, H=w, D=24.0, r_min=2.91, mean_radius=1.2, size_sigma=0.25, morph=0,
                        seed=seed)


def _field(h, w, ppp):
    yy, xx = np.mgrid[0:h, 0:w]
    rel = 0.05 + 2.95 * (0.5 + 0.5 * np.sin(xx / 9.0) * np.cos(yy / 13.0))
    return (rel * 1.0417 / ppp).astype(np.float32)


def _cfg(E, w, ppp, seed=11, **kw):
    d = dict(W=float(w), H=float(w), D=24.0, g=0.85, r_ahu=0.14, mu_abs=0.9, theta_max=0.12,
             max_bounces=48, ppp=ppp, seed=seed, fe=1.2, sigma=0.1041666667, w_px=w, h_px=w,
             spp=4, ss=1, scan_seed=12, **TRIX_OPTICS)
    d.update(kw)
    return O.expected_cfg(E.shape, **d)


@pytest.mark.parametrize("thr", [0, 1, 2, 4, 9])
def test_poisson_tail_against_scipy(thr):
    lam = np.array([0.0, 1e-6, 0.01, 0.5, 3.9, 4.0, 4.1, 12.0, 60.0, 1e4])
    p, dp = O.poisson_tail(lam, thr)
    ref = poisson.sf(thr - 1, lam) if thr > 0 else np.ones_like(lam)
    ok = ref > 1e-300
    assert np.allclose(p[ok], ref[ok], rtol=1e-10, atol=1e-300)
    dref = poisson.pmf(thr - 1, lam) if thr > 0 else np.zeros_like(lam)
    assert np.allclose(dp, dref, rtol=1e-10, atol=1e-300)


def test_hg_score_quadrature_identities():
    """E[score] = 0 and E[mu * score] = d E[mu] / dg = 1 under the HG density."""
    for g in (0.3, 0.85, -0.4):
        pdf = lambda m: 0.5 * (1 - g * g) / (1 + g * g - 2 * g * m) ** 1.5
        m0, _ = integrate.quad(lambda m: pdf(m) * O.hg_score(g, m), -1, 1, limit=400, points=[1.0])
        m1, _ = integrate.quad(lambda m: pdf(m) * m * O.hg_score(g, m), -1, 1, limit=400, points=[1.0])
        assert abs(m0) < 1e-8 and m1 == pytest.approx(1.0, rel=1e-7)


def test_single_cube_dose_and_absorption_tangent_closed_form():
    """Collimated normal beam, no scattering, no Fresnel loss: every packet inside the
    footprint deposits 5 E s (1 - exp(-2 a mu)); d dose / d mu over dose is
    2 a exp(-2 a mu) / (1 - exp(-2 a mu)); reflectance 0 deposits nothing more."""
    a, mu, s, ppp = 1.5, 0.4, 1.3, 64
    cs = _one([10, 10, 5], a, s)
    E = np.full((20, 20), 0.8, np.float32)
    cfg = O.expected_cfg(E.shape, W=20.0, H=20.0, D=10.0, mu_a=1e-12, mu_s=0.0, g=0.0, n_refr=1.0,
                         r_ahu=0.0, od_ahu=0.0, mu_abs=mu, theta_max=0.0, max_bounces=4, ppp=ppp,
                         seed=5, fe=1.0, sigma=0.1, w_px=20, h_px=20)
    f = O.expected_forward(cs, E, cfg, 4, 3)
    per = 5 * 0.8 * s * (1 - np.exp(-2 * a * mu))
    n_hit = f["dose"][0] / per
    assert abs(n_hit - round(n_hit)) < 1e-7 * n_hit  # an integer number of packets crossed
    N = 400 * ppp
    p = (2 * a) ** 2 / 400.0
    assert abs(n_hit - N * p) < 4 * np.sqrt(N * p * (1 - p))
    v = O.expected_vjp(cs, E, cfg, 4, 3, np.zeros((20, 20)), f["dose"])
    gm = np.zeros(E.shape)
    gr = np.zeros(3)
    import ctypes as C
    cc = cs.c()
    one = np.ones(1)
    O.lib().orc_replay_grad(C.byref(cc), O._p(E, O._f32p), C.byref(cfg), O._p(one, O._f64p),
                            O._p(gm, O._f64p), O._p(gr, O._f64p))
    assert gr[0] / f["dose"][0] == pytest.approx(2 * a * np.exp(-2 * a * mu) / (1 - np.exp(-2 * a * mu)),
                                                 rel=1e-9)
    assert gr[1] == 0.0                                 # no scattering: no score
    assert (gm * E).sum() == pytest.approx(f["dose"][0], rel=1e-12)   # Euler: linear in E
    assert (v["grad_exposure"] == 0).all()


def test_expected_scan_single_cube_closed_form():
    """T = 1 - p (1 - exp(-sigma 2 a fe b)) inside the boosted footprint, 1 outside;
    dT/dp = -(1 - exp(-sigma 2 a fe b)) and dT/db = -p sigma 2 a fe exp(-sigma 2 a fe b)
    for a pixel fully inside."""
    a, fe, sig, p = 2.0, 1.2, 0.3, 0.37
    cs = _one([10, 10, 5], a)
    E = np.ones((20, 20), np.float32)
    cfg = O.expected_cfg(E.shape, W=20.0, H=20.0, D=10.0, mu_a=0.0, mu_s=0.0, g=0.0, n_refr=1.0,
                         r_ahu=0.0, od_ahu=0.0, mu_abs=1.0, theta_max=0.0, max_bounces=1, ppp=1,
                         seed=1, fe=fe, sigma=sig, w_px=20, h_px=20, spp=4)
    for b in (1.0, 1.3):
        T = O.expected_scan(cs, [p], [b], cfg)
        L = 2 * a * fe * b
        assert T[10, 10] == pytest.approx(1 - p * (1 - np.exp(-sig * L)), rel=1e-12) and T[0, 0] == 1.0
        gT = np.zeros((20, 20))
        gT[10, 10] = 1.0
        gp, gb = O.expected_scan(cs, [p], [b], cfg, gT)
        assert gp[0] == pytest.approx(-(1 - np.exp(-sig * L)), rel=1e-12)
        assert gb[0] == pytest.approx(-p * sig * 2 * a * fe * np.exp(-sig * L), rel=1e-12)
    # one crystal has no neighbours: no boost
    assert O.touch_graph(cs, fe, 1.25).nnz == 0


def test_scan_adjoint_matches_finite_differences_of_prob_and_boost():
    cs = _tile(32)
    E = _field(32, 32, 8)
    cfg = _cfg(E, 32, 8)
    rng = np.random.default_rng(1)
    prob = rng.uniform(0.05, 0.95, cs.n)
    boost = rng.uniform(1.0, 2.0, cs.n)
    gT = rng.standard_normal((32, 32))
    gp, gb = O.expected_scan(cs, prob, boost, cfg, gT)
    d = rng.standard_normal(cs.n)
    h = 1e-6
    fd = ((O.expected_scan(cs, prob + h * d, boost, cfg) - O.expected_scan(cs, prob - h * d, boost, cfg)) * gT).sum() / (2 * h)
    assert fd == pytest.approx(gp @ d, rel=1e-6)
    fd = ((O.expected_scan(cs, prob, boost + h * d, cfg) - O.expected_scan(cs, prob, boost - h * d, cfg)) * gT).sum() / (2 * h)
    assert fd == pytest.approx(gb @ d, rel=1e-5)


@pytest.mark.parametrize("morph", [0, 1])
def test_chord_scale_derivative_against_finite_differences(morph):
    """d chord / d s of a rotated crystal scaled about its centre, for rays from outside
    and from inside."""
    rng = np.random.default_rng(7 + morph)
    n = 40
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    dims = rng.uniform(0.5, 2.0, (n, 3))
    cs = O.Crystals(np.zeros((n, 3)), dims, q, np.full(n, morph, np.int32), np.ones(n))
    checked = 0
    for i in range(n):
        for inside in (False, True):
            d = rng.standard_normal(3)
            d /= np.linalg.norm(d)
            o = (rng.uniform(-0.3, 0.3, 3) if inside else -6.0 * d + rng.uniform(-0.6, 0.6, 3))
            s = 1.3
            r = O.chord_scaled(cs, i, s, o, d)
            rp, rm = O.chord_scaled(cs, i, s + 1e-6, o, d), O.chord_scaled(cs, i, s - 1e-6, o, d)
            if r is None or rp is None or rm is None:
                continue
            assert r[1] == pytest.approx((rp[0] - rm[0]) / 2e-6, rel=1e-5, abs=1e-7)
            checked += 1
    assert checked > 40


def test_touch_graph_against_brute_force():
    cs = _tile(28)
    fe, pf = 1.32, 1.35
    A = O.touch_graph(cs, fe, pf).toarray() > 0
    m = cs.dims.max(axis=1)
    dist = np.linalg.norm(cs.pos[:, None, :] - cs.pos[None, :, :], axis=2)
    ref = dist <= (m[:, None] + m[None, :]) * fe * pf
    np.fill_diagonal(ref, False)
    assert (A == ref).all() and (A == A.T).all() and A.sum() > cs.n


def test_clump_boost_follows_the_reference_boost_at_integer_sizes():
    """development.rs:164-171: 1.15 + 0.18 (min(s, 6) - 1) for a clump of s >= 2."""
    s = np.array([2, 3, 4, 5, 6, 7, 20], float)
    b, _ = O.clump_boost(s)
    assert np.allclose(b, 1.15 + 0.18 * (np.minimum(s, 6) - 1))
    b, db = O.clump_boost(np.array([0.0, 1.0, 1.5, 2.0 - 1e-12, 2.0 + 1e-12]))
    assert np.allclose(b, [1.0, 1.0, 1.165, 1.33, 1.33])
    x = np.linspace(1.01, 7.9, 300)
    bb, dd = O.clump_boost(x)
    fd = (O.clump_boost(x + 1e-7)[0] - O.clump_boost(x - 1e-7)[0]) / 2e-7
    ok = (np.abs(x - 2) > 1e-3) & (np.abs(x - 6) > 1e-3)
    assert np.allclose(fd[ok], dd[ok], atol=1e-6)


@pytest.fixture(scope="module")
def small():
    w, ppp = 40, 32
    cs = _tile(w)
    E = _field(w, w, ppp)
    tgt = O.expected_forward(cs, E * np.float32(2 ** 0.3), _cfg(E, w, ppp, seed=99), 4, 3)["out"]
    return cs, E, w, ppp, tgt


def _loss_grad(cs, E, cfg, tgt):
    f = O.expected_forward(cs, E, cfg, 4, 3)
    r = f["out"] - tgt
    return float(np.mean(r * r)), 2 * r / r.size, f


def test_pathwise_gradients_match_central_differences(small):
    """mu_abs, R and a random exposure direction: central differences with common random
    numbers (h = 1e-3 relative) agree with the replayed adjoint."""
    cs, E, w, ppp, tgt = small
    cfg = _cfg(E, w, ppp)
    L0, G, f = _loss_grad(cs, E, cfg, tgt)
    v = O.expected_vjp(cs, E, cfg, 4, 3, G, f["dose"])
    for k, (key, val) in enumerate([("mu_abs", 0.9), (None, None), ("r_ahu", 0.14)]):
        if key is None:
            continue
        h = 1e-3 * val
        Lp = _loss_grad(cs, E, _cfg(E, w, ppp, **{key: val + h}), tgt)[0]
        Lm = _loss_grad(cs, E, _cfg(E, w, ppp, **{key: val - h}), tgt)[0]
        assert (Lp - Lm) / (2 * h) == pytest.approx(v["grads"][k], rel=1e-3)
    d = np.random.default_rng(2).standard_normal(E.shape)
    Ep = (E * (1 + 1e-3 * d)).astype(np.float32)
    Em = (E * (1 - 1e-3 * d)).astype(np.float32)
    fd = _loss_grad(cs, Ep, cfg, tgt)[0] - _loss_grad(cs, Em, cfg, tgt)[0]
    adj = (v["grad_exposure"] * (Ep.astype(np.float64) - Em)).sum()
    assert fd == pytest.approx(adj, rel=1e-3)


def test_dose_is_homogeneous_in_exposure(small):
    """Euler: sum_px E dL/dE = sum_i dL/ddose_i dose_i for the linear dose map."""
    cs, E, w, ppp, tgt = small
    cfg = _cfg(E, w, ppp)
    L0, G, f = _loss_grad(cs, E, cfg, tgt)
    v = O.expected_vjp(cs, E, cfg, 4, 3, G, f["dose"])
    assert (v["grad_exposure"] * E).sum() == pytest.approx(v["grad_dose"] @ f["dose"], rel=1e-9)


def test_reflectance_derivative_at_zero_is_first_order_term(small):
    """At R = 0 the reflected dose is R times the once-reflected dose plus O(R^2)."""
    cs, E, w, ppp, tgt = small
    cfg0 = _cfg(E, w, ppp, r_ahu=0.0)
    L0, G, f = _loss_grad(cs, E, cfg0, tgt)
    v = O.expected_vjp(cs, E, cfg0, 4, 3, G, f["dose"])
    h = 1e-4
    Lh = _loss_grad(cs, E, _cfg(E, w, ppp, r_ahu=h), tgt)[0]
    assert (Lh - L0) / h == pytest.approx(v["grads"][2], rel=2e-3)
    assert v["grads"][2] != 0.0


def test_hg_score_gradient_agrees_with_finite_differences_on_average():
    """g enters through the scattering directions, so central differences of one sample
    jump where paths cross the slab faces; over seeds the score-function gradient and
    h = 3e-2 central differences estimate the same derivative (within their standard
    errors, measured over the seeds).  The score estimate differentiates the loss of the
    expected dose; finite differences, the expected loss of the sampled dose: the two
    differ by O(1 / packets per crystal), so the packet count is kept high."""
    w, ppp = 48, 256
    cs = _tile(w)
    E = _field(w, w, ppp)
    tgt = O.expected_forward(cs, E * np.float32(2 ** 0.3), _cfg(E, w, ppp, seed=99), 4, 3)["out"]
    sc, fd = [], []
    for seed in range(21, 27):
        cfg = _cfg(E, w, ppp, seed=seed)
        L0, G, f = _loss_grad(cs, E, cfg, tgt)
        sc.append(O.expected_vjp(cs, E, cfg, 4, 3, G, f["dose"])["grads"][1])
        h = 0.85 * 3e-2
        fd.append((_loss_grad(cs, E, _cfg(E, w, ppp, seed=seed, g=0.85 + h), tgt)[0] -
                   _loss_grad(cs, E, _cfg(E, w, ppp, seed=seed, g=0.85 - h), tgt)[0]) / (2 * h))
    sc, fd = np.array(sc), np.array(fd)
    se = np.sqrt(sc.var(ddof=1) / len(sc) + fd.var(ddof=1) / len(fd))
    assert abs(sc.mean() - fd.mean()) < 3.5 * se
    assert abs(sc.mean() - fd.mean()) < 0.2 * abs(sc.mean())
    # the score estimate is the low-variance one
    assert sc.std() < fd.std()
