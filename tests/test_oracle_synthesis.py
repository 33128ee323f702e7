"""Pins of the device-synthesis twin in the oracle (dart throwing; no reference
counterpart) -- run on the CPU.  The reference's own checks of its sampler
(poisson.rs:141-165 minimum distance, emulsion.rs:233-252 bounds and sizes) apply to any
Poisson-disk sampler; saturation and the attribute distributions are checked against
closed forms."""
import numpy as np
import pytest
from scipy import stats
from scipy.spatial import cKDTree

from oracle import oracle as O

KW = dict(W=48.0, H=40.0, D=24.0, r_min=2.91, mean_radius=1.2, size_sigma=0.25, seed=5)


@pytest.mark.parametrize("morph", [0, 1, 2])
def test_darts_minimum_distance_bounds_and_determinism(morph):
    cs = O.synthesize_darts(morph=morph, aspect=8.0 if morph == 1 else 1.0, **KW)
    p = cs.pos
    assert cs.n > 100
    assert (p >= 0).all() and (p[:, 0] < 48).all() and (p[:, 1] < 40).all() and (p[:, 2] < 24).all()
    d = cKDTree(p).query(p, k=2)[0][:, 1]
    assert d.min() >= 2.91 * (1 - 1e-6)
    again = O.synthesize_darts(morph=morph, aspect=8.0 if morph == 1 else 1.0, **KW)
    assert np.array_equal(again.pos, p) and np.array_equal(again.dims, cs.dims)
    other = O.synthesize_darts(morph=morph, aspect=8.0 if morph == 1 else 1.0, **dict(KW, seed=6))
    assert not np.array_equal(other.pos[:10], p[:10])


def test_darts_saturate_the_slab():
    """After the default 30 rounds almost no point of the slab is farther than r from
    every centre (a Poisson-disk set is maximal when none is); the density is near the
    random-sequential-addition jamming density n r^3 = 0.73."""
    cs = O.synthesize_darts(morph=0, **KW)
    rng = np.random.default_rng(0)
    probe = rng.uniform([0, 0, 0], [48, 40, 24], (20000, 3))
    free = cKDTree(cs.pos).query(probe)[0] >= 2.91
    assert free.mean() < 0.02
    dens = cs.n * 2.91 ** 3 / (48 * 40 * 24)
    assert 0.66 < dens < 0.80


def test_darts_attributes_follow_the_reference_distributions():
    """emulsion.rs:160-207: radius LogNormal(ln mean, sigma) clamped to
    [min_radius, 0.48 r_min], sensitivity N(base, 0.15) clamped at 0.1, unit quaternions,
    tabular half-thickness r / aspect with tilt within 0.17 rad."""
    cs = O.synthesize_darts(morph=0, **dict(KW, W=96.0, H=96.0))
    r = cs.dims[:, 0]
    assert np.allclose(cs.dims, r[:, None])
    inside = (r > 0.15) & (r < 0.48 * 2.91)
    # conditional KS on the unclamped range
    lo, hi = stats.norm.cdf(np.log([0.15, 0.48 * 2.91]), np.log(1.2), 0.25)
    u = (stats.norm.cdf(np.log(r[inside]), np.log(1.2), 0.25) - lo) / (hi - lo)
    assert stats.kstest(u, "uniform").pvalue > 0.01
    assert stats.kstest(cs.sens, stats.norm(1.2, 0.15).cdf).pvalue > 0.01
    assert np.allclose(np.linalg.norm(cs.quat, axis=1), 1.0)
    t = O.synthesize_darts(morph=1, aspect=8.0, **KW)
    assert np.allclose(t.dims[:, 2], np.maximum(t.dims[:, 0] / 8.0, 0.05))
    tilt = 2 * np.arccos(np.clip(t.quat[:, 0], -1, 1))
    assert tilt.max() <= 0.17 + 1e-9 and abs(t.quat[:, 3]).max() == 0.0
