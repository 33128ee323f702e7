"""Pins of the oracle (the CPU restatement) -- run on the CPU.

Each check is chosen so that a plausible restatement mistake fails it: the
reference's own golden values and invariants (file:line cited), textbook closed forms,
and independent re-derivations (scipy rotations, scipy connected components, a
recursive visit order) rather than a retyped formula.
"""
import numpy as np
import pytest
from scipy import integrate
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components
from scipy.spatial.transform import Rotation

from oracle import oracle as O


def crystals(pos, dims, quat=None, morph=0, sens=1.0, developed=None, fe=None):
    pos = np.asarray(pos, float).reshape(-1, 3)
    n = len(pos)
    dims = np.broadcast_to(np.asarray(dims, float), (n, 3))
    quat = np.broadcast_to(np.asarray(quat if quat is not None else [1, 0, 0, 0], float), (n, 4))
    return O.Crystals(pos, dims, quat, np.broadcast_to(morph, (n,)), np.broadcast_to(sens, (n,)),
                      developed, fe)


# ------------------------------------------------------------------ random streams
def test_philox_known_answer_vectors():
    """Random123 known-answer vectors for Philox4x32-10."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, exp in kat:
        assert list(O.philox4x32_10(ctr, key)) == exp


# ------------------------------------------------------------------ crystal geometry
@pytest.mark.parametrize("fp32", [False, True])
def test_cubic_crystal_known_answer(fp32):
    """crystal.rs:238-261: unit cube at (10,10,10), ray from z=0 along +z -> (9, 11)."""
    cs = crystals([[10, 10, 10]], [1, 1, 1])
    t = O.crystal_intersect(cs, 0, [10, 10, 0], [0, 0, 1], fp32)
    assert abs(t[0] - 9) < 1e-4 and abs(t[1] - 11) < 1e-4
    assert O.crystal_intersect(cs, 0, [20, 10, 0], [0, 0, 1], fp32) is None


@pytest.mark.parametrize("fp32", [False, True])
def test_tabular_crystal_known_answer(fp32):
    """crystal.rs:264-282: ellipsoid (2,2,0.2) at z=5 -> (4.8, 5.2)."""
    cs = crystals([[0, 0, 5]], [2, 2, 0.2], morph=1)
    t = O.crystal_intersect(cs, 0, [0, 0, 0], [0, 0, 1], fp32)
    assert abs(t[0] - 4.8) < 1e-4 and abs(t[1] - 5.2) < 1e-4


def _slab_local(lo, ld, h):
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / ld
        a = (-h - lo) * inv
        b = (h - lo) * inv
    tmin = np.fmax.reduce(np.fmin(a, b))
    tmax = np.fmin.reduce(np.fmax(a, b))
    return (max(tmin, 0.0), tmax) if (tmax >= tmin and tmax > 0) else None


def test_rotated_crystals_against_scipy_rotation():
    """Oriented box and ellipsoid hits against an independent derivation: scipy's
    rotation of the world ray into the crystal frame (orientation maps local->world)."""
    rng = np.random.default_rng(0)
    q = Rotation.random(300, random_state=1).as_quat()  # (x, y, z, w)
    quat = np.concatenate([q[:, 3:], q[:, :3]], 1)
    pos = rng.uniform(-1, 1, (300, 3))
    dims = rng.uniform(0.5, 2.0, (300, 3))
    for morph in (0, 1):
        cs = O.Crystals(pos, dims, quat, np.full(300, morph), np.ones(300))
        hits = 0
        for i in range(300):
            o = rng.uniform(-4, 4, 3)
            d = pos[i] + rng.uniform(-1.5, 1.5, 3) - o  # aim near the crystal
            d /= np.linalg.norm(d)
            R = Rotation.from_quat(q[i])
            lo, ld = R.inv().apply(o - pos[i]), R.inv().apply(d)
            if morph == 0:
                exp = _slab_local(lo, ld, dims[i])
            else:
                p, dd = lo / dims[i], ld / dims[i]
                A, B, Cc = dd @ dd, 2 * p @ dd, p @ p - 1
                disc = B * B - 4 * A * Cc
                exp = None
                if disc >= 0:
                    r0, r1 = (-B - np.sqrt(disc)) / (2 * A), (-B + np.sqrt(disc)) / (2 * A)
                    exp = (max(min(r0, r1), 0.0), max(r0, r1)) if max(r0, r1) > 0 else None
            got = O.crystal_intersect(cs, i, o, d)
            assert (got is None) == (exp is None), (morph, i)
            if got is not None:
                hits += 1
                assert got[0] == pytest.approx(exp[0], abs=1e-9)
                assert got[1] == pytest.approx(exp[1], abs=1e-9)
        assert hits > 100


def test_sphere_chord_closed_form():
    """An ellipsoid with equal semi-axes r is a sphere: chord 2 sqrt(r^2 - b^2)."""
    cs = crystals([[0, 0, 10]], [1.5, 1.5, 1.5], morph=2)
    for b in np.linspace(0, 1.4, 8):
        t0, t1 = O.crystal_intersect(cs, 0, [b, 0, 0], [0, 0, 1])
        assert t1 - t0 == pytest.approx(2 * np.sqrt(1.5 ** 2 - b ** 2), rel=1e-12)


def test_expansion_scales_developed_crystals_only():
    """crystal.rs:145-156: developed crystals use dims * filament_expansion."""
    base = crystals([[0, 0, 10]] * 2, [1, 1, 1], developed=np.array([1, 0], np.uint8),
                    fe=np.array([1.5, 1.5]))
    t0, t1 = O.crystal_intersect(base, 0, [0, 0, 0], [0, 0, 1])
    assert t1 - t0 == pytest.approx(3.0)
    t0, t1 = O.crystal_intersect(base, 1, [0, 0, 0], [0, 0, 1])
    assert t1 - t0 == pytest.approx(2.0)


# ------------------------------------------------------------------------------ BVH
def test_bvh_construction_and_query_known_answer():
    """bvh.rs:412-439: 20 unit cubes at x = 3i; ray at x=15 along +z hits only 5."""
    cs = crystals([[3 * i, 0, 10] for i in range(20)], [1, 1, 1])
    nh, idx, _, _ = O.all_hits(cs, [[15, 0, 0]], [[0, 0, 1]])
    assert nh[0] == 1 and idx[0, 0] == 5


def test_bvh_radius_query_known_answer():
    """bvh.rs:442-462: crystals at x = i, radius 1.5 about x=5 -> [4, 5, 6]."""
    cs = crystals([[i, 0, 0] for i in range(10)], [0.4, 0.4, 0.4])
    assert sorted(O.query_radius(cs, [5, 0, 0], 1.5).tolist()) == [4, 5, 6]


def _random_scene(n, seed, morph=0, size=(40, 40, 12)):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 1, (n, 3)) * np.array(size)
    q = Rotation.random(n, random_state=seed).as_quat()
    quat = np.concatenate([q[:, 3:], q[:, :3]], 1)
    dims = np.repeat(rng.uniform(0.3, 1.2, (n, 1)), 3, 1)
    return O.Crystals(pos, dims, quat, np.full(n, morph), np.ones(n))


def test_bvh_structure_invariants():
    cs = _random_scene(500, 1)
    b = O.bvh_build(cs)
    nn = len(b["left"])
    assert sorted(b["indices"].tolist()) == list(range(500))
    m = cs.dims.max(1, keepdims=True)
    amin, amax = cs.pos - m, cs.pos + m
    for k in range(nn):
        if b["count"][k] > 0:
            assert b["count"][k] <= 4
            ids = b["indices"][b["offset"][k]:b["offset"][k] + b["count"][k]]
            assert np.array_equal(b["bmin"][k], amin[ids].min(0))
            assert np.array_equal(b["bmax"][k], amax[ids].max(0))
        else:
            assert b["left"][k] == k + 1  # interior reserved before its children
            for c in (b["left"][k], b["right"][k]):
                assert (b["bmin"][c] >= b["bmin"][k]).all() and (b["bmax"][c] <= b["bmax"][k]).all()


def test_all_hits_order_is_right_first_dfs_and_complete_for_axis_aligned_cubes():
    """For axis-aligned cubes (no corner leaves its box) the BVH hit set equals the
    brute-force set; the order equals a recursive visit that descends right before
    left, the order the reference's stack produces (bvh.rs:339-343)."""
    rng = np.random.default_rng(3)
    n = 400
    pos = rng.uniform(0, 1, (n, 3)) * np.array([30, 30, 10])
    cs = O.Crystals(pos, np.full((n, 3), 0.8), np.tile([1.0, 0, 0, 0], (n, 1)), np.zeros(n),
                    np.ones(n))
    b = O.bvh_build(cs)

    def box_hit(k, o, inv):
        with np.errstate(invalid="ignore"):
            t1 = (b["bmin"][k] - o) * inv
            t2 = (b["bmax"][k] - o) * inv
        te = np.fmax.reduce(np.fmin(t1, t2))
        tx = np.fmin.reduce(np.fmax(t1, t2))
        return tx >= te and tx > 0

    for r in range(200):
        o = np.array([rng.uniform(0, 30), rng.uniform(0, 30), -1.0])
        d = rng.normal(size=3)
        d[2] = abs(d[2]) + 0.3
        d /= np.linalg.norm(d)
        with np.errstate(divide="ignore"):
            inv = 1.0 / d
        order = []

        def visit(k):
            if not box_hit(k, o, inv):
                return
            if b["count"][k] > 0:
                for s in range(b["offset"][k], b["offset"][k] + b["count"][k]):
                    i = b["indices"][s]
                    if O.crystal_intersect(cs, int(i), o, d) is not None:
                        order.append(int(i))
            else:
                visit(b["right"][k])
                visit(b["left"][k])

        visit(0)
        nh, idx, _, _ = O.all_hits(cs, [o], [d])
        assert idx[0, :nh[0]].tolist() == order
        brute = [i for i in range(n) if O.crystal_intersect(cs, i, o, d) is not None]
        assert sorted(order) == brute


def test_first_hit_is_the_smallest_entry_over_all_crystals():
    """The oracle's first hit (the twin of fg_traverse_first_hit): for axis-aligned cubes,
    where the BVH finds every crystal a ray crosses, it is the brute-force smallest
    t_enter over all crystals (f32 twin), ties to the earlier hit in the DFS order; for
    rotated cubes it is the first smallest entry of intersect_ray_all's list."""
    rng = np.random.default_rng(11)
    n = 500
    pos = rng.uniform(0, 1, (n, 3)) * np.array([30, 30, 10])
    aligned = O.Crystals(pos, np.full((n, 3), 0.8), np.tile([1.0, 0, 0, 0], (n, 1)), np.zeros(n), np.ones(n))
    q = Rotation.random(n, random_state=4).as_quat()[:, [3, 0, 1, 2]]
    rotated = O.Crystals(pos, np.full((n, 3), 0.8), q, np.zeros(n), np.ones(n))
    o = np.stack([rng.uniform(0, 30, 300), rng.uniform(0, 30, 300), rng.uniform(-1, 10, 300)], 1)
    d = rng.normal(size=(300, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o, d = o.astype(np.float32), d.astype(np.float32)
    for cs, brute in ((aligned, True), (rotated, False)):
        idx, t0, _ = O.first_hit(cs, o, d, threads=3)
        nh, oi, ot0, _ = O.all_hits(cs, o, d, max_hits=256, fp32=True)
        assert (idx >= 0).sum() > 100
        for r in range(len(o)):
            if nh[r] == 0:
                assert idx[r] == -1
                continue
            j = int(np.argmin(ot0[r, :nh[r]]))
            assert idx[r] == oi[r, j] and t0[r] == ot0[r, j]
            if brute:
                hits = [(O.crystal_intersect(cs, i, o[r], d[r], fp32=True), i) for i in range(n)]
                best = min(h[0][0] for h in hits if h[0] is not None)
                tied = [i for h, i in hits if h is not None and h[0] == best]
                dfs = [int(i) for i in oi[r, :nh[r]]]
                assert idx[r] == min(tied, key=dfs.index)


def test_halation_branch_against_its_closed_form():
    """The anti-halation step of the transport (transport.rs:156-163, optics.rs:160-196)
    against a closed form: a collimated normal beam, no gelatin scattering or absorption,
    one wide thin crystal near the top, reflectance R with OD 0 (helpers.halation_expected:
    the downward chord, then the reflected weight back along the same line with
    probability 0.6 or cosine-diffusely with the chord cut by the step cap).  A branch
    that skips the backing factor, reflects downward, scatters after the reflection or
    swaps the diffuse and specular probabilities misses it by tens of standard errors."""
    from helpers import halation_expected, halation_scene
    sc = halation_scene()
    cs = O.Crystals(np.array(sc["pos"]), np.array(sc["dims"]), np.array(sc["quat"]), np.zeros(1), np.ones(1))
    for R in (0.0, 1.0):
        means = []
        for seed in range(4):
            c, st = O.expose(cs, None, W=sc["W"], H=sc["W"], D=sc["D"], mu_a=0.0, mu_s=0.0, g=0.85, n_refr=1.5,
                             r_ahu=R, od_ahu=0.0, mu_abs=2.0, theta_max=0.0, max_bounces=48, seed=100 + seed,
                             sampler=0, photons_per_area=256.0, threads=4)
            means.append(int(c.sum()) / st["launched"])
        m, se = np.mean(means), np.std(means, ddof=1) / 2.0
        exp = halation_expected(2.0, sc["t"], sc["zc"], sc["D"], 1.5, R)
        assert abs(m - exp) < 4.5 * se + 1e-4 * exp, (R, m, exp, se)
        if R > 0:
            wrong = halation_expected(2.0, sc["t"], sc["zc"], sc["D"], 1.5, R, p_diffuse=0.6)
            assert abs(wrong - exp) > 20 * se
            no_backing = halation_expected(2.0, sc["t"], sc["zc"], sc["D"], 1.5, 0.0)
            assert abs(no_backing - exp) > 100 * se


# ---------------------------------------------------------------------- development
def test_development_known_answer():
    """development.rs:197-265."""
    cs = crystals([[0, 0, 5], [0.8, 0, 5], [1.6, 0, 5], [10, 0, 5], [20, 0, 5]], [0.5, 0.5, 0.5])
    dev, fe, cid, rep = O.develop(cs, np.array([10, 10, 10, 4, 1], np.uint32), 3, 2.0, 1.5)
    assert rep["developed"] == 4 and rep["fixed"] == 1
    assert dev[4] == 0
    assert cid[0] == cid[1] == cid[2] and cid[3] != cid[0]
    assert rep["clumps"] == 2 and rep["max_clump"] == 3
    # clump boost 1.15 + 0.18 * (min(3, 6) - 1) on the 3-clump, none on the singleton
    assert fe[:3] == pytest.approx(2.0 * (1.15 + 0.18 * 2))
    assert fe[3] == 2.0 and fe[4] == 0.0


def test_clumps_are_connected_components():
    rng = np.random.default_rng(5)
    n = 600
    cs = crystals(rng.uniform(0, 1, (n, 3)) * np.array([30, 30, 8]), [0.0, 0.0, 0.0])
    cs.dims = np.repeat(rng.uniform(0.2, 0.7, (n, 1)), 3, 1)
    counts = rng.integers(0, 8, n).astype(np.uint32)
    fe0, pf, thr = 1.3, 1.2, 3
    dev, fe, cid, rep = O.develop(cs, counts, thr, fe0, pf)
    dv = np.nonzero(counts >= thr)[0]
    assert (dev.astype(bool) == (counts >= thr)).all()
    m = cs.dims.max(1) * fe0
    P = cs.pos[dv]
    d2 = ((P[:, None, :] - P[None, :, :]) ** 2).sum(-1)
    touch = (m[dv][:, None] + m[dv][None, :]) * pf
    ii, jj = np.nonzero(d2 <= touch ** 2)
    ncomp, lab = connected_components(coo_matrix((np.ones(len(ii)), (ii, jj)), shape=(len(dv),) * 2),
                                      directed=False)
    assert rep["clumps"] == ncomp
    # same partition, numbered by first member
    first = {}
    for k, l in enumerate(lab):
        first.setdefault(l, len(first))
    assert [first[l] for l in lab] == cid[dv].tolist()
    sizes = np.bincount(lab)
    boost = np.where(sizes[lab] > 1, 1.15 + 0.18 * (np.minimum(sizes[lab], 6) - 1), 1.0)
    assert np.allclose(fe[dv], fe0 * boost, rtol=0, atol=1e-15)
    assert rep["max_clump"] == sizes.max()


# ----------------------------------------------------------------------- scanning
def test_transmittance_known_answer_and_closed_form():
    """transmittance.rs:181-220 plus T = exp(-sigma * 2 h) inside an axis-aligned
    cube's footprint and 1 outside."""
    cs = crystals([[5, 5, 4]], [2, 2, 1], developed=np.ones(1, np.uint8), fe=np.ones(1))
    T = O.scan(cs, 10, 10, 8, 16, 16, spp=4, sigma=20.0, seed=42)
    assert T[8, 8] < T[0, 0] and T[0, 0] > 0.95
    T = O.scan(cs, 10, 10, 8, 16, 16, spp=4, sigma=0.3, seed=1)
    assert T[8, 8] == pytest.approx(np.exp(-0.3 * 2.0), rel=1e-6)
    assert T[0, 0] == 1.0


def test_scan_empty_and_dneg():
    cs = crystals([[5, 5, 4]], [1, 1, 1], developed=np.zeros(1, np.uint8))
    T = O.scan(cs, 10, 10, 8, 8, 8, spp=4, sigma=1.0, seed=0)
    assert (T == 1.0).all()
    D = O.dneg(np.array([1.0, 1e-5, 0.1], np.float32))
    assert D[0] == pytest.approx(0.05) and D[1] == pytest.approx(4.05) and D[2] == pytest.approx(1.05)


def test_scan_supersample_uniform_field_invariant():
    """A developed slab covering the whole film gives the same T for any aperture."""
    cs = crystals([[5, 5, 4]], [50, 50, 1], developed=np.ones(1, np.uint8), fe=np.ones(1))
    T1 = O.scan(cs, 10, 10, 8, 8, 8, spp=4, ss=1, sigma=0.5, seed=3)
    T2 = O.scan(cs, 10, 10, 8, 8, 8, spp=4, ss=2, sigma=0.5, seed=3)
    assert np.allclose(T1, np.exp(-1.0), rtol=1e-6) and np.allclose(T2, T1, rtol=1e-6)


def test_otf_against_independent_numpy():
    rng = np.random.default_rng(2)
    img = rng.random((13, 17)).astype(np.float32)
    k = np.array([0.035, 0.93, 0.035], np.float32)
    p = np.pad(img, ((0, 0), (1, 1)), mode="edge")
    h = (p[:, :-2] * k[0] + p[:, 1:-1] * k[1]) + p[:, 2:] * k[2]
    p = np.pad(h, ((1, 1), (0, 0)), mode="edge")
    v = (p[:-2] * k[0] + p[1:-1] * k[1]) + p[2:] * k[2]
    assert (O.otf3(img).view(np.uint32) == v.astype(np.float32).view(np.uint32)).all()
    c = np.full((5, 6), 0.7, np.float32)
    assert np.allclose(O.otf3(c), 0.7, atol=1e-6)


# ------------------------------------------------------------------------ optics
def test_free_path_mean_closed_form():
    """optics.rs:206-224: mean free path 1/mu_t (quadrature over the uniform)."""
    mu = 0.05
    m, _ = integrate.quad(lambda u: O.free_path(mu, u), 0, 1, limit=200)
    assert m == pytest.approx(1 / mu, rel=1e-6)
    assert O.free_path(0.0, 0.5) == 1e6


@pytest.mark.parametrize("g", [0.0, 0.5, 0.85, -0.3])
def test_hg_mean_cosine_equals_g(g):
    d = np.array([0.3, -0.2, 0.9])
    d /= np.linalg.norm(d)
    xs = (np.arange(4000) + 0.5) / 4000
    cos = np.array([O.hg(d, g, x, 0.37) @ d for x in xs])
    assert cos.mean() == pytest.approx(g, abs=2e-4)
    out = O.hg(d, 0.85, 0.3, 0.8)
    assert np.linalg.norm(out) == pytest.approx(1.0)


def test_fresnel_schlick():
    r0 = ((1 - 1.534) / (1 + 1.534)) ** 2
    assert O.fresnel(1.534, 1.0) == pytest.approx(r0)
    assert O.fresnel(1.534, 0.0) == pytest.approx(1.0)
    assert O.fresnel(1.534, -1.0) == pytest.approx(r0)


def test_antihalation_known_answer():
    """optics.rs:227-245: energy decreases, reflected direction points up (-z)."""
    for seed in range(20):
        alive, e, d = O.antihalation(0.12, 0.85, 999 + seed, 1.0, [0, 0, 1])
        if alive:
            assert e < 1.0 and d[2] < 0 and np.linalg.norm(d) == pytest.approx(1.0)


# ---------------------------------------------------------------------- transport
def test_single_crystal_collimated_closed_form():
    """One axis-aligned cube (semi-axis a) under a collimated normal beam with no
    scattering and no Fresnel loss: E[counts] = N (4a^2 / WH) 5 E s (1 - exp(-2 a mu))."""
    a, mu, E, s = 1.5, 0.4, 0.8, 1.3
    cs = crystals([[10, 10, 5]], [a, a, a], sens=s)
    emap = np.full((4, 4), E, np.float32)
    counts, st = O.expose(cs, emap, W=20, H=20, D=10, mu_a=1e-12, mu_s=0.0, g=0.0, n_refr=1.0,
                          r_ahu=0.0, od_ahu=0.0, mu_abs=mu, theta_max=1e-9, max_bounces=4, seed=9,
                          sampler=1, n_photons=400000)
    N = st["launched"]
    p = 4 * a * a / 400.0
    lam = 5 * E * s * (1 - np.exp(-2 * a * mu))
    exp = N * p * lam
    # binomial spread of the number of packets crossing, plus the rounding Bernoullis
    sd = np.sqrt(N * p * (1 - p)) * lam + np.sqrt(N * p * 0.25)
    assert abs(counts[0] - exp) < 4 * sd
    assert st["fresnel_rejected"] == 0


def test_empty_slab_deposits_nothing():
    cs = crystals([[10, 10, 50]], [1, 1, 1])  # outside the slab
    counts, st = O.expose(cs, None, W=20, H=20, D=10, mu_a=0.003, mu_s=0.05, g=0.8, n_refr=1.5,
                          r_ahu=0.1, od_ahu=0.5, mu_abs=1.0, theta_max=0.1, max_bounces=48, seed=1,
                          sampler=1, n_photons=20000)
    assert counts.sum() == 0 and st["launched"] == 20000


def test_reference_chunking_rounds_launch_up():
    """transport.rs:65-72: chunks = 8 * threads, ceil(base / chunks) packets per chunk."""
    cs = crystals([[10, 10, 5]], [1, 1, 1])
    _, st = O.expose(cs, None, W=7, H=3, D=10, mu_a=0.003, mu_s=0.05, g=0.8, n_refr=1.5,
                     r_ahu=0.1, od_ahu=0.5, mu_abs=1.0, theta_max=0.1, max_bounces=48, seed=1,
                     sampler=0, photons_per_area=10.0, threads=3)
    base = 210
    assert st["chunks"] == 24 and st["launched"] == 24 * -(-base // 24)


def test_exposure_latent_image_formation_oracle():
    """transport.rs:196-237 on the oracle: left half lit -> left absorbs > 5x right."""
    cs = O.synthesize(W=20, H=20, D=14, r_min=1.2, mean_radius=0.55, size_sigma=0.4, morph=0,
                      base_sens=1.2, seed=100)
    emap = np.zeros((2, 2), np.float32)
    emap[:, 0] = 1.0
    counts, _ = O.expose(cs, emap, W=20, H=20, D=14, mu_a=0.003, mu_s=0.05, g=0.8, n_refr=1.534,
                         r_ahu=0.14, od_ahu=0.8, mu_abs=1.5, theta_max=0.1, max_bounces=32, seed=42,
                         sampler=0, photons_per_area=100.0, threads=4)
    x = cs.pos[:, 0]
    left, right = counts[x <= 9].sum(), counts[x >= 11].sum()
    assert left > 0 and left > 5 * right


# ---------------------------------------------------------------------- synthesis
def test_poisson_min_distance_and_bounds():
    """poisson.rs:141-165 and emulsion.rs:233-252."""
    cs = O.synthesize(W=20, H=20, D=10, r_min=2.0, mean_radius=0.5, size_sigma=0.3, morph=0, seed=42)
    d2 = ((cs.pos[:, None] - cs.pos[None]) ** 2).sum(-1) + np.eye(cs.n) * 1e9
    assert d2.min() >= (2.0 - 1e-6) ** 2
    assert (cs.pos >= 0).all() and (cs.pos <= [20, 20, 10]).all()
    assert (cs.dims >= 0.15).all() and (cs.dims <= 0.96 + 1e-12).all()
    assert (cs.sens >= 0.1).all()
    assert np.allclose(np.linalg.norm(cs.quat, axis=1), 1.0)


def test_tabular_synthesis_tilt_and_thickness():
    cs = O.synthesize(W=20, H=20, D=8, r_min=0.9, mean_radius=0.4, size_sigma=0.2, morph=1,
                      aspect=6.5, seed=3)
    assert (cs.dims[:, 2] == np.maximum(cs.dims[:, 0] / 6.5, 0.05)).all()
    angle = 2 * np.arccos(np.clip(cs.quat[:, 0], -1, 1))
    assert (angle <= 0.17 + 1e-12).all()
    assert np.allclose(cs.quat[:, 3], 0.0)  # tilt axis lies in the film plane


# ------------------------------------------------------------ scan high-pass (delta_d)
def test_highpass17_against_scipy_correlation():
    """gaussian.wgsl:21-59 twice (gpu.rs:1180-1206): the oracle's f32 passes against
    scipy's separable correlation with edge replication in f64; a tap shifted by one
    column or a pass on the wrong axis fails by orders of magnitude more."""
    from scipy.ndimage import correlate1d
    rng = np.random.default_rng(11)
    d = rng.uniform(0.05, 3.0, (37, 53)).astype(np.float32)
    w = O.GAUSS17.astype(np.float64)
    want = d - correlate1d(correlate1d(d.astype(np.float64), w, axis=1, mode="nearest"), w, axis=0,
                           mode="nearest")
    got = O.highpass17(d)
    assert np.abs(got - want).max() < 2e-6
    shifted = d - correlate1d(correlate1d(d.astype(np.float64), w, axis=1, mode="nearest", origin=1), w,
                              axis=0, mode="nearest")
    assert np.abs(shifted - want).max() > 1e-2


def test_highpass17_closed_forms():
    """Flat field c: c (1 - S^2) with S = 1.146 the sum of the shader's weights; an
    impulse far from the edges: delta minus the outer product of the weights; a field of
    one row or one column is blurred along one axis only, the other axis repeating it."""
    S = float(O.GAUSS17.astype(np.float64).sum())
    assert S == pytest.approx(1.146, abs=1e-7)   # f32 weights
    flat = O.highpass17(np.full((20, 30), 0.7, np.float32))
    assert np.allclose(flat, 0.7 * (1.0 - S * S), atol=1e-6)
    imp = np.zeros((41, 41), np.float32)
    imp[20, 20] = 1.0
    w = O.GAUSS17.astype(np.float64)
    want = -np.outer(w, w)
    want[8, 8] += 1.0
    assert np.abs(O.highpass17(imp)[12:29, 12:29] - want).max() < 1e-7
    row = np.random.default_rng(3).uniform(0, 1, (1, 25)).astype(np.float32)
    from scipy.ndimage import correlate1d
    assert np.abs(O.highpass17(row) - (row - S * correlate1d(row.astype(np.float64), w, axis=1,
                                                              mode="nearest"))).max() < 1e-6
