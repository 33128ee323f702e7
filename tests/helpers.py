"""Shared builders for tests: small emulsions in pixel units and the oracle view of a
product emulsion (same crystals)."""
import numpy as np

import fgrain_amd as F
from fgrain_amd import workloads as W
from oracle import oracle as O

TRIX = W.load_config("tile320_trix_cubic")


def trix_config(w, h, depth=24.0, g=0.85, r_ahu=0.0, morph=0, aspect=1.0, mean=1.2, r_min=2.91):
    cfg = W.emulsion_config(TRIX, float(w), float(h))
    cfg.dimensions = (float(w), float(h), depth)
    cfg.optics.anisotropy_g = g
    cfg.optics.antihalation_reflectance = r_ahu
    cfg.morphology = F.Morphology(morph)
    cfg.tabular_aspect_ratio = aspect
    cfg.mean_crystal_radius = mean
    cfg.r_min = r_min
    return cfg


def oracle_crystals(em: F.VirtualEmulsionLayer, developed=None, fe=None) -> O.Crystals:
    c = em.crystals()
    return O.Crystals(c["pos"], c["dims"], c["quat"], c["morph"], c["sens"], developed, fe)


def expose_kwargs(cfg: F.EmulsionConfig, mu_abs=0.9, theta=0.12, max_bounces=48):
    o = cfg.optics
    W_, H_, D_ = cfg.dimensions
    return dict(W=W_, H=H_, D=D_, mu_a=o.absorption_coeff, mu_s=o.scattering_coeff,
                g=o.anisotropy_g, n_refr=o.refractive_index, r_ahu=o.antihalation_reflectance,
                od_ahu=o.antihalation_density, mu_abs=mu_abs, theta_max=theta,
                max_bounces=max_bounces)


def random_rays(n, W_, H_, D_, seed, scattered_frac=0.5):
    """Origins on the top face or inside the slab; directions in the 0.12 rad cone or
    uniform on the sphere."""
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(0, W_, n), rng.uniform(0, H_, n),
                  np.where(rng.random(n) < 0.5, 0.001, rng.uniform(0, D_, n))], 1)
    th = rng.uniform(0, 0.12, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    cone = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    d = np.where((rng.random(n) < scattered_frac)[:, None], v, cone)
    return o.astype(np.float32), d.astype(np.float32)


def halation_scene(W=64.0, D=24.0, zc=2.0, t=1.0):
    """One cubic crystal wider than the patch (semi-axes 1000, 1000, t/2) centred at depth
    zc: every packet of a collimated normal beam crosses it on the way down and again on a
    specular reflection; a diffuse reflection crosses it when its step reaches it."""
    return dict(pos=[[W / 2, W / 2, zc]], dims=[[1000.0, 1000.0, t / 2]], quat=[[1.0, 0, 0, 0]], W=W, D=D,
                zc=zc, t=t)


def halation_expected(mu, t, zc, D, n_refr, R, od=0.0, sens=1.0, p_diffuse=0.4):
    """Expected latent photons per launched packet in halation_scene under a collimated
    normal beam with no gelatin scattering or absorption (transport.rs:89-184,
    optics.rs:147-196): Fresnel rejection r0 at normal incidence; the way down deposits
    5 s pa (pa = 1 - exp(-mu t)) and leaves weight max(1 - 0.9 pa, 0.05); the backing
    scales it by R 10^(-2 od); with probability 1 - p_diffuse the packet returns along the
    same line (chord t again), otherwise in a cosine-weighted direction c = cos(theta),
    whose chord through the crystal is cut by the step cap 1.5 D (the packet restarts at
    D - 0.001); a scattering event ends its deposits (albedo 0)."""
    from scipy import integrate
    r0 = ((1 - n_refr) / (1 + n_refr)) ** 2
    pa = 1 - np.exp(-mu * t)
    rho = max(1 - 0.9 * pa, 0.05) * R * 10 ** (-2 * od)
    z0 = D - 0.001
    a, b, cap = z0 - (zc + t / 2), z0 - (zc - t / 2), 1.5 * D

    def f(c):
        L = max(min(b / c, cap) - a / c, 0.0)
        return (1 - np.exp(-mu * L)) * 2 * c

    brk = sorted({a / cap, b / cap})
    diff, _ = integrate.quad(f, 0.0, 1.0, points=brk, epsabs=1e-13, epsrel=1e-12, limit=200)
    return (1 - r0) * 5 * sens * (pa + rho * ((1 - p_diffuse) * pa + p_diffuse * diff))
