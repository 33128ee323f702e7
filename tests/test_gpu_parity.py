"""GPU parity of the hot path against the oracle (CPU restatement of the reference).

Every test calls through the C ABI (libfgrain_hip.so) and compares with the oracle on
the same seeded inputs:
* traversal: all-hit lists and first hits bit-exact against the FP32 oracle;
* transport: same Philox streams replayed by the FP64 oracle (near-identical counts),
  and an independent-stream statistical comparison with the reference's chunked
  uniform sampler;
* development: exact (integer threshold, clump partition, expansion);
* scan: |T_gpu - T_oracle| <= 1e-5 on identical rays and developed sets.
"""
import numpy as np
import pytest
from scipy import stats

import fgrain_amd as F
from fgrain_amd import workloads as W
from helpers import expose_kwargs, oracle_crystals, random_rays, trix_config
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_em():
    cfg = trix_config(96, 96)
    return cfg, F.VirtualEmulsionLayer.synthesize(cfg, 7)


@pytest.mark.parametrize("morph", [0, 1, 2])
def test_all_hits_bit_exact_vs_fp32_oracle(engine, morph):
    cfg = trix_config(64, 64, morph=morph, aspect=8.0 if morph == 1 else 1.0)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 11 + morph)
    o, d = random_rays(4096, 64, 64, 24, seed=morph)
    nh, idx, t0, t1 = engine.traverse_all_hits(em, o, d, max_hits=128)
    cs = oracle_crystals(em)
    onh, oidx, ot0, ot1 = O.all_hits(cs, o, d, max_hits=128, fp32=True)
    assert (nh == onh).all()
    assert nh.sum() > 1000
    for r in range(len(nh)):
        k = nh[r]
        assert (idx[r, :k] == oidx[r, :k]).all()
        assert (t0[r, :k].view(np.uint32) == ot0[r, :k].view(np.uint32)).all()
        assert (t1[r, :k].view(np.uint32) == ot1[r, :k].view(np.uint32)).all()


def test_first_hit_bit_exact_and_fp64_mismatch_rate(engine, small_em):
    cfg, em = small_em
    o, d = random_rays(1 << 16, 96, 96, 24, seed=3)
    idx, t0, t1 = engine.first_hit(em, o, d)
    cs = oracle_crystals(em)
    onh, oidx, ot0, ot1 = O.all_hits(cs, o, d, max_hits=256, fp32=True)
    exp = np.full(len(o), -1, np.int32)
    e0 = np.zeros(len(o), np.float32)
    for r in np.nonzero(onh)[0]:
        k = onh[r]
        j = int(np.argmin(ot0[r, :k]))  # argmin returns the first minimum: DFS tie order
        exp[r], e0[r] = oidx[r, j], ot0[r, j]
    assert (idx == exp).all()
    hit = idx >= 0
    assert (t0[hit].view(np.uint32) == e0[hit].view(np.uint32)).all()
    # informational bound: the float path agrees with the double reference on
    # nearly every ray
    dnh, didx, dt0, _ = O.all_hits(cs, o.astype(np.float64), d.astype(np.float64), max_hits=256)
    d_exp = np.full(len(o), -1, np.int32)
    for r in np.nonzero(dnh)[0]:
        d_exp[r] = didx[r, int(np.argmin(dt0[r, :dnh[r]]))]
    assert np.mean(d_exp != idx) < 1e-3


def _expose_gpu(engine, em, cfg, emap, ppp, seed, **kw):
    x = F.ExposureConfig(0.0, kw.get("theta", 0.12), kw.get("mu_abs", 0.9),
                         kw.get("max_bounces", 48), ppp)
    st = F.expose_emulsion(em, emap, x, seed, engine)
    return em.absorbed_photons().astype(np.int64), st


def test_transport_replay_matches_fp64_oracle(engine):
    """Same photon-keyed Philox streams, f32 GPU vs f64 oracle: paths only diverge
    where rounding flips a comparison, so counts agree almost everywhere."""
    cfg = trix_config(64, 64)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 5)
    emap = np.linspace(0.01, 4.0, 64, dtype=np.float32)[None, :].repeat(64, 0)
    g, st = _expose_gpu(engine, em, cfg, emap, ppp=16, seed=77)
    cs = oracle_crystals(em)
    oc, ost = O.expose(cs, emap, **expose_kwargs(cfg), seed=77, sampler=1, ppp=16)
    oc = oc.astype(np.int64)
    assert st["launched"] == ost["launched"] == 16 * 64 * 64
    assert ost["abs_total"] > 0
    assert abs(g.sum() - oc.sum()) <= 1e-3 * oc.sum()
    assert np.mean(g == oc) >= 0.98
    # the kernel stops a packet once its weight is exactly zero (halation off zeroes it;
    # it can deposit nothing afterwards): compare the work of packets with weight
    assert st["bounces"] == pytest.approx(ost["bounces_live"], rel=1e-3)
    # the kernel skips subtrees that cannot hold an absorbing crystal (exact pruning);
    # the reference walks the full infinite ray
    assert 0.2 * ost["nodes_live"] < st["nodes_visited"] <= ost["nodes_live"] * 1.001


@pytest.mark.parametrize("tuning", [dict(transport_kernel=1), dict(transport_wait_pct=100),
                                    dict(transport_wait_pct=1, transport_chunk=640)])
def test_transport_kernels_and_tunings_give_identical_counts(engine, tuning):
    """The persistent lane-regenerating kernel and the one-thread-per-packet kernel, under
    any tuning, deposit exactly the same latent photons (same draws, same hit order)."""
    cfg = trix_config(48, 40, r_ahu=0.6)
    cfg.optics.antihalation_density = 0.1
    emap = np.linspace(0.01, 4.0, 48, dtype=np.float32)[None, :].repeat(40, 0)
    ref = F.VirtualEmulsionLayer.synthesize(cfg, 12)
    engine.set_tuning()
    _, st0 = _expose_gpu(engine, ref, cfg, emap, ppp=24, seed=3)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 12)
    try:
        engine.set_tuning(**tuning)
        _, st1 = _expose_gpu(engine, em, cfg, emap, ppp=24, seed=3)
    finally:
        engine.set_tuning()
    assert np.array_equal(ref.absorbed_photons(), em.absorbed_photons())
    assert st0["bounces"] == st1["bounces"] and st0["abs_total"] == st1["abs_total"]
    assert st0["prims_tested"] == st1["prims_tested"]


def test_transport_steps_that_overflow_the_leaf_queue(engine):
    """A 200-deep layer with little scattering (mean free path ~330 px, step cap 300):
    one step crosses the whole layer and the 40-odd leaf boxes on its way, several times
    the persistent kernel's per-lane leaf queue, so lanes stop with a full queue, test
    it and resume their walk.  The one-thread-per-packet kernel tests each leaf when it
    is hit (no queue): both must deposit exactly the same photons, and agree with the
    oracle's replay of the same streams."""
    cfg = trix_config(32, 32, depth=200.0)
    cfg.optics.scattering_coeff = 0.001
    emap = np.full((32, 32), 1.0, np.float32)
    ref = F.VirtualEmulsionLayer.synthesize(cfg, 21)
    engine.set_tuning()
    g, st0 = _expose_gpu(engine, ref, cfg, emap, ppp=8, seed=4)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 21)
    try:
        engine.set_tuning(transport_kernel=1)
        _, st1 = _expose_gpu(engine, em, cfg, emap, ppp=8, seed=4)
    finally:
        engine.set_tuning()
    assert np.array_equal(ref.absorbed_photons(), em.absorbed_photons())
    assert st0["prims_tested"] == st1["prims_tested"] and st0["abs_total"] == st1["abs_total"]
    # crystals tested per bounce (about 100 in the oracle's full walk): a leaf holds at
    # most 4 crystals here, so more than 48 means more leaves than one 12-deep queue
    assert st0["prims_tested"] > 48 * st0["bounces"]
    cs = oracle_crystals(ref)
    oc, ost = O.expose(cs, emap, **expose_kwargs(cfg), seed=4, sampler=1, ppp=8)
    oc = oc.astype(np.int64)
    assert ost["abs_total"] > 0
    assert abs(g.sum() - oc.sum()) <= 1e-3 * oc.sum()
    assert np.mean(g == oc) >= 0.98


def test_transport_replay_with_halation(engine):
    cfg = trix_config(48, 48, depth=16.0, r_ahu=0.9)
    cfg.optics.antihalation_density = 0.05
    em = F.VirtualEmulsionLayer.synthesize(cfg, 9)
    emap = np.full((48, 48), 1.5, np.float32)
    g, st = _expose_gpu(engine, em, cfg, emap, ppp=32, seed=5)
    cs = oracle_crystals(em)
    oc, ost = O.expose(cs, emap, **expose_kwargs(cfg), seed=5, sampler=1, ppp=32)
    assert abs(g.sum() - oc.astype(np.int64).sum()) <= 1e-3 * oc.sum()
    assert np.mean(g == oc) >= 0.97


def test_transport_replay_tabular_with_film_base(engine):
    """Tabular (ellipsoid) crystals, halation on below a film base (the lateral offset
    extension, also in the oracle): replayed streams agree."""
    cfg = trix_config(64, 64, depth=19.2, r_ahu=0.6, morph=1, aspect=8.0, mean=0.9, r_min=2.16)
    cfg.optics.antihalation_density = 0.0
    cfg.min_half_thickness = 0.05
    em = F.VirtualEmulsionLayer.synthesize(cfg, 31)
    emap = np.full((64, 64), 2.0, np.float32)
    x = F.ExposureConfig(0.0, 0.12, 0.9, 48, 16, base_thickness=10.0)
    st = F.expose_emulsion(em, emap, x, 8, engine)
    g = em.absorbed_photons().astype(np.int64)
    cs = oracle_crystals(em)
    oc, ost = O.expose(cs, emap, **expose_kwargs(cfg), seed=8, sampler=1, ppp=16, base_thickness=10.0)
    assert ost["abs_total"] > 0
    oc = oc.astype(np.int64)
    assert abs(g.sum() - oc.sum()) <= 1e-3 * oc.sum()
    # reflected packets live long and the base offset divides by the direction's z, so
    # float and double paths drift apart more often than without the base; a wrong sign
    # or scale of the offset would move most crystals by far more than this
    assert np.mean(g == oc) >= 0.85
    assert np.abs(g - oc).sum() <= 0.01 * oc.sum()
    assert st["bounces"] == pytest.approx(ost["bounces_live"], rel=2e-3)
    # the base offset changes the result (it is not silently ignored)
    em2 = F.VirtualEmulsionLayer.synthesize(cfg, 31)
    F.expose_emulsion(em2, emap, F.ExposureConfig(0.0, 0.12, 0.9, 48, 16), 8, engine)
    assert not np.array_equal(em2.absorbed_photons(), em.absorbed_photons())


def test_transport_statistics_vs_reference_sampler(engine):
    """Independent streams: GPU (stratified photon-keyed Philox) against the reference
    algorithm's chunked uniform sampler (transport.rs:65-74).  Per-crystal means agree
    in aggregate and in bins, and the per-crystal distributions pass a two-sample KS."""
    cfg = trix_config(96, 96)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 21)
    emap = np.linspace(0.01, 4.0, 96, dtype=np.float32)[None, :].repeat(96, 0)
    ppp = 64
    cum = []
    for s in range(4):
        c, _ = _expose_gpu(engine, em, cfg, emap, ppp=ppp, seed=1000 + s)
        cum.append(c)
    # absorbed_photons accumulates over calls (record_absorption adds): difference them
    cum = np.stack(cum)
    per = np.diff(np.concatenate([np.zeros((1, cum.shape[1]), np.int64), cum]), axis=0)
    gmean = per.mean(0)
    cs = oracle_crystals(em)
    n = ppp * 96 * 96
    oc, ost = O.expose(cs, emap, **expose_kwargs(cfg), seed=3, sampler=0,
                       photons_per_area=ppp, threads=8)
    # the reference launches chunks * ceil(base / chunks) >= base packets
    oc = oc.astype(np.float64) * (n / ost["launched"])
    tot_g, tot_o = gmean.sum(), oc.sum()
    sd = per.sum(1).std(ddof=1)
    assert abs(tot_g - tot_o) <= 4 * sd * np.sqrt(1 + 1 / 4) + 1e-3 * tot_o
    assert abs(tot_g - tot_o) / tot_o < 5e-3
    pos = cs.pos
    bins = np.minimum((pos[:, 0] / 96 * 8).astype(int), 7)
    for b in range(8):
        m = bins == b
        assert abs(gmean[m].sum() - oc[m].sum()) / oc[m].sum() < 0.02
    assert stats.ks_2samp(per[0], oc).pvalue > 0.01


def test_transport_means_at_1e8_photons_vs_reference_sampler(engine):
    """The stated bar at full scale on the C1 tile (320 x 320 x 24 px Tri-X, ramp
    0.01..4.0, mu_abs 0.9, g 0.85, halation off): 1.05e8 GPU packets (16 runs of 64 per
    pixel) against 1e8 packets of the reference's chunked uniform sampler on the host
    (transport.rs:65-74).  Aggregate mean within 0.1 %, x-bins within 0.5 %, per-crystal
    z-scores centred with unit spread (each side's variance measured from 16 replicate
    runs of its sampler), two-sample KS p > 0.01."""
    import os
    from fgrain_amd import workloads as W
    cfg = W.load_config("tile320_trix_cubic")
    em = F.VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, 320.0, 320.0), cfg["seed"])
    emap = W.workload_exposure(cfg)
    x = W.exposure_config(cfg)
    runs = []
    zero = np.zeros(len(em), np.uint32)
    for k in range(16):
        em.set_state(counts=zero)
        F.expose_emulsion(em, emap, x, 7000 + k, engine)
        runs.append(em.absorbed_photons().astype(np.float64))
    runs = np.stack(runs)
    n_gpu = 16 * 64 * 320 * 320
    g = runs.sum(0) / n_gpu                               # latent counts per packet
    var_g = runs.var(0, ddof=1) * 16 / n_gpu ** 2         # variance of g
    e, o = cfg["emulsion"], cfg["emulsion"]["optics"]
    cs = oracle_crystals(em)
    oc, ost = O.expose(cs, emap, W=320.0, H=320.0, D=e["depth"], mu_a=o["absorption_coeff"],
                       mu_s=o["scattering_coeff"], g=o["anisotropy_g"], n_refr=o["refractive_index"],
                       r_ahu=o["antihalation_reflectance"], od_ahu=o["antihalation_density"],
                       mu_abs=x.crystal_absorption_coeff, theta_max=x.illumination_angle_rad,
                       max_bounces=x.max_bounces, seed=11, sampler=0,
                       photons_per_area=1e8 / (320.0 * 320.0), threads=min(os.cpu_count() or 1, 16))
    n_cpu = ost["launched"]
    c = oc.astype(np.float64) / n_cpu
    # variance of the uniform-position sampler per packet, from GPU runs in that mode
    # (photons_per_pixel 0: uniform positions over the patch, like the reference)
    xu = F.ExposureConfig(64.0, x.illumination_angle_rad, x.crystal_absorption_coeff, x.max_bounces, 0)
    ur = []
    for k in range(16):
        em.set_state(counts=zero)
        F.expose_emulsion(em, emap, xu, 9000 + k, engine)
        ur.append(em.absorbed_photons().astype(np.float64))
    n_u = 64 * 320 * 320
    var_c = np.stack(ur).var(0, ddof=1) / n_u / n_cpu
    print("aggregate rel diff %.2e" % (g.sum() / c.sum() - 1))
    assert abs(g.sum() / c.sum() - 1) < 1e-3
    bins = np.minimum((cs.pos[:, 0] / 320 * 8).astype(int), 7)
    for b in range(8):
        m = bins == b
        assert abs(g[m].sum() / c[m].sum() - 1) < 5e-3
    ok = (var_g + var_c) > 0
    z = (g[ok] - c[ok]) / np.sqrt(var_g[ok] + var_c[ok])
    print("z mean %.3f sd %.3f over %d crystals" % (z.mean(), z.std(), ok.sum()))
    assert abs(z.mean()) < 5 / np.sqrt(ok.sum()) + 0.02
    assert 0.85 < z.std() < 1.15
    assert stats.ks_2samp(g * n_gpu / 16, c * n_gpu / 16).pvalue > 0.01


def test_halation_branch_against_its_closed_form(engine):
    """The device transport's anti-halation step (transport.rs:156-163, optics.rs:160-196)
    against the closed form of helpers.halation_expected (collimated normal beam, no
    gelatin scattering, one wide thin crystal, reflectance R with OD 0): 16 seeds of 2^22
    packets each; the standard error comes from the spread of the seeds.  Swapping the
    diffuse and specular probabilities, or dropping the backing, misses by far more."""
    from helpers import halation_expected, halation_scene
    sc = halation_scene()
    for R in (0.0, 1.0):
        cfg = trix_config(sc["W"], sc["W"], depth=sc["D"], r_ahu=R)
        cfg.optics.absorption_coeff = 0.0
        cfg.optics.scattering_coeff = 0.0
        cfg.optics.refractive_index = 1.5
        cfg.optics.antihalation_density = 0.0
        means = []
        for seed in range(16):
            em = F.VirtualEmulsionLayer.from_crystals(cfg, sc["pos"], sc["dims"], sc["quat"], 0, 1.0)
            st = F.expose_emulsion(em, None, F.ExposureConfig(1024.0, 0.0, 2.0, 48, 0), 300 + seed, engine)
            assert st["launched"] == 1 << 22
            means.append(int(em.absorbed_photons().sum()) / st["launched"])
        m, se = np.mean(means), np.std(means, ddof=1) / 4.0
        exp = halation_expected(2.0, sc["t"], sc["zc"], sc["D"], 1.5, R)
        print(f"R {R}: mean {m:.6f} +- {se:.2e}, closed form {exp:.6f}", flush=True)
        assert abs(m - exp) < 4.5 * se + 1e-4 * exp
        if R > 0:
            assert abs(halation_expected(2.0, sc["t"], sc["zc"], sc["D"], 1.5, R, p_diffuse=0.6) - exp) > 20 * se


def test_c1_tile_first_hits_bit_exact(engine):
    """The C1 tile (tile320_trix_cubic: 320 x 320 x 24 px, the reference seed's Tri-X
    emulsion): first hits of 2^20 seeded rays (half from the top face, half inside the
    slab; half in the 0.12 rad illumination cone, half in all directions) equal the FP32
    oracle's bit for bit (index, t_enter, t_exit)."""
    cfg = W.load_config("tile320_trix_cubic")
    em = F.VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, 320.0, 320.0), cfg["seed"])
    assert len(em) > 55000
    o, d = random_rays(1 << 20, 320, 320, 24, seed=20)
    idx, t0, t1 = engine.first_hit(em, o, d)
    cs = oracle_crystals(em)
    oi, ot0, ot1 = O.first_hit(cs, o, d)
    assert (oi >= 0).sum() > (1 << 19)
    assert (idx == oi).all()
    hit = oi >= 0
    assert (t0[hit].view(np.uint32) == ot0[hit].view(np.uint32)).all()
    assert (t1[hit].view(np.uint32) == ot1[hit].view(np.uint32)).all()


@pytest.mark.parametrize("ss,scale", [(1, 1.0), (2, 1.0)])
def test_c1_tile_scan_matches_fp64_oracle(engine, ss, scale):
    """The C1 tile exposed, developed and scanned as the workload defines it (ramp 0.01-4.0,
    64 photons per pixel, threshold 4, fe 1.2, pf 1.25, 4 rays per aperture; at that
    exposure nearly every crystal develops and T stays below 0.04): the scan of the
    developed set (transmittance.rs:65-171) at supersample 1 and 2 is within 1e-5 of the
    FP64 oracle on the same rays, and D_neg within the same bound carried through the log."""
    cfg = W.load_config("tile320_trix_cubic")
    em = F.VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, 320.0, 320.0), cfg["seed"])
    emap = W.workload_exposure(cfg) * np.float32(scale)
    F.expose_emulsion(em, emap, W.exposure_config(cfg), cfg["seed"] + 1, engine)
    rep = F.develop_emulsion(em, W.development_config(cfg), engine)
    assert 0 < rep.developed_crystals < len(em)
    st = em.state()
    s = cfg["scan"]
    sc = F.ScanConfig(320, 320, s["samples_per_pixel"], F.ScanMode.NegativeTransmittance, s["silver_opacity"], ss)
    T, _ = F.scan_emulsion(em, sc, cfg["seed"] + 2, engine)
    cs = oracle_crystals(em, st["developed"], st["fe"])
    To = O.scan(cs, 320, 320, 24, 320, 320, spp=s["samples_per_pixel"], ss=ss, sigma=s["silver_opacity"],
                seed=cfg["seed"] + 2)
    print(f"C1 scan ss {ss} x{scale}: {rep.developed_crystals} of {len(em)} developed, "
          f"T {To.min():.4f}..{To.max():.4f}, max |dT| {np.abs(T - To).max():.2e}", flush=True)
    if scale < 1:
        assert To.min() < 0.1 and To.max() > 0.9
    assert np.abs(T - To).max() <= 1e-5
    sc.mode = F.ScanMode.NegativeDensity
    D, _ = F.scan_emulsion(em, sc, cfg["seed"] + 2, engine)
    # D_neg = -log10(clamp(T, 1e-4, 1)) + 0.05: the T tolerance carried through the log
    bound = 1e-5 / (np.log(10.0) * np.maximum(To, 1e-4)) + 1e-5
    assert (np.abs(D - O.dneg(To)) <= bound).all()


def test_development_exact(engine):
    """develop_emulsion with an engine runs the device kernels (threshold, union-find
    clumping over the tile's BVH, boost; development.rs:59-188): exact against the
    oracle."""
    cfg = trix_config(96, 96)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 4)
    rng = np.random.default_rng(0)
    counts = rng.poisson(3.0, len(em)).astype(np.uint32)
    em.set_state(counts=counts)
    rep = F.develop_emulsion(em, F.DevelopmentConfig(4, 1.2, 1.25), engine)
    st = em.state()
    cs = oracle_crystals(em)
    dev, fe, cid, orep = O.develop(cs, counts, 4, 1.2, 1.25)
    assert (st["developed"] == dev).all()
    assert (st["clump_id"] == cid).all()
    assert (st["fe"] == fe).all()
    assert rep.developed_crystals == orep["developed"]
    assert rep.total_clumps == orep["clumps"]
    assert rep.max_clump_size == orep["max_clump"]
    assert rep.avg_clump_size == pytest.approx(orep["avg_clump"], rel=1e-12)


@pytest.mark.parametrize("morph,ss", [(0, 1), (0, 2), (1, 1), (1, 2), (2, 1)])
def test_scan_matches_fp64_oracle(engine, morph, ss):
    """Every morphology (cubes, 8:1 tabular platelets, octahedral ellipsoids): the scan
    solves each crystal from the level of its centre, so thin platelets far below the
    ray origin keep their chords in f32 (from the origin, b^2 - 4ac cancels to 1e-3 of
    the chord)."""
    cfg = trix_config(64, 48, morph=morph, aspect=8.0 if morph == 1 else 1.0)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 8)
    rng = np.random.default_rng(1)
    counts = rng.poisson(3.2, len(em)).astype(np.uint32)
    em.set_state(counts=counts)
    F.develop_emulsion(em, F.DevelopmentConfig(4, 1.2, 1.25), engine)
    st = em.state()
    sc = F.ScanConfig(64, 48, 4, F.ScanMode.NegativeTransmittance, 0.104166667, ss)
    T, _ = F.scan_emulsion(em, sc, 2, engine)
    cs = oracle_crystals(em, st["developed"], st["fe"])
    To = O.scan(cs, 64, 48, 24, 64, 48, spp=4, ss=ss, sigma=0.104166667, seed=2)
    assert T.min() < 0.9
    assert np.abs(T - To).max() <= 1e-5
    sc.mode = F.ScanMode.NegativeDensity
    D, _ = F.scan_emulsion(em, sc, 2, engine)
    assert np.abs(D - O.dneg(To)).max() <= 1e-4


def test_scan_refills_its_leaf_queue_and_matches_oracle(engine):
    """A layer 200 deep: every vertical ray crosses more than 16 leaf boxes (the scan's
    per-lane leaf queue depth), so each ray fills the queue and resumes its walk several
    times; the result must still equal the oracle's depth-first sum."""
    cfg = trix_config(24, 24, depth=200.0)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 5)
    # leaf boxes of the full layer that contain each pixel centre: with every crystal
    # developed the scan's tree splits the same centroids with boxes at least as large,
    # so this counts a lower bound of the leaves each ray visits
    nodes, _ = em.device_layout()
    f = nodes.view(np.float32)
    leaf = nodes[:, 7] != 0xFFFFFFFF
    lo, hi = f[leaf, 0:2], f[leaf, 3:5]
    xs = np.arange(24) + 0.5
    per_ray = np.array([[np.sum((lo[:, 0] < x) & (x < hi[:, 0]) & (lo[:, 1] < y) & (y < hi[:, 1]))
                         for x in xs] for y in xs])
    assert per_ray.min() > 16 and per_ray.max() > 48
    em.set_state(counts=np.full(len(em), 50, np.uint32))
    F.develop_emulsion(em, F.DevelopmentConfig(4, 1.2, 1.25), engine)
    st = em.state()
    assert st["developed"].all()
    sc = F.ScanConfig(24, 24, 4, F.ScanMode.NegativeTransmittance, 0.001, 1)
    T, _ = F.scan_emulsion(em, sc, 7, engine)
    cs = oracle_crystals(em, st["developed"], st["fe"])
    To = O.scan(cs, 24, 24, 200, 24, 24, spp=4, ss=1, sigma=0.001, seed=7)
    assert 0.2 < To.min() and To.max() < 0.8
    assert np.abs(T - To).max() <= 1e-5


def test_scan_otf_matches_oracle(engine):
    cfg = trix_config(40, 40)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 3)
    em.set_state(counts=np.full(len(em), 9, np.uint32))
    F.develop_emulsion(em, F.DevelopmentConfig(4, 1.2, 1.25), engine)
    sc = F.ScanConfig(40, 40, 4, F.ScanMode.NegativeTransmittance, 0.104166667, 1, False)
    T, _ = F.scan_emulsion(em, sc, 2, engine)
    sc.otf = True
    To, _ = F.scan_emulsion(em, sc, 2, engine)
    ref = O.otf3(T)
    assert (To.view(np.uint32) == ref.view(np.uint32)).all()


def test_reference_exposure_latent_image_formation(engine):
    """transport.rs:196-237: Tri-X 20x20 um, left half lit -> left absorbs > 5x right."""
    cfg = F.EmulsionConfig.tri_x_400(20.0, 20.0)
    em = F.VirtualEmulsionLayer.synthesize(cfg, 100)
    emap = np.zeros((64, 64), np.float32)
    emap[:, :32] = 1.0  # nearest-pixel map of the closure x <= 10
    F.expose_emulsion(em, emap, F.ExposureConfig(100.0, 0.1, 1.5, 32), 42, engine)
    c = em.absorbed_photons()
    x = em.crystals()["pos"][:, 0]
    left, right = c[x <= 9.0].sum(), c[x >= 11.0].sum()
    assert left > 0 and left > 5 * right


def test_reference_transmittance_evaluation(engine):
    """transmittance.rs:181-220: one developed 2x2x1 um cube at (5,5,4): centre pixel
    darker than the corner, corner above 0.95."""
    cfg = F.EmulsionConfig.t_max_100(10.0, 10.0)
    em = F.VirtualEmulsionLayer.from_crystals(cfg, [[5, 5, 4]], [[2, 2, 1]], [[1, 0, 0, 0]], 0, 1.0)
    em.set_state(developed=np.ones(1, np.uint8), fe=np.ones(1))
    T, _ = F.scan_emulsion(em, F.ScanConfig(16, 16, 4, F.ScanMode.NegativeTransmittance, 20.0), 42,
                           engine)
    assert T[8, 8] < T[0, 0]
    assert T[0, 0] > 0.95
    # closed form: inside the footprint T = exp(-sigma * 2 * 1.0)
    assert T[8, 8] == pytest.approx(np.exp(-20.0 * 2.0), abs=1e-6)


@pytest.mark.parametrize("exp_val", [0.0, 0.05, 0.18, 0.5, 1.0])
def test_reference_render_tile_all_gpu(engine, exp_val):
    """gpu.rs:1280-1317 on the drop-in: a 16x16 tile renders 256 finite D_neg values."""
    p = F.GpuTileParams.for_preset("tri-x", 16 / 15.0, 16, 16, exposure=1.0, tile_seed=42)
    out = engine.render_tile_all_gpu(np.full(256, exp_val, np.float32), p)
    assert out.shape == (256,)
    assert np.isfinite(out).all()
    assert (out >= 0.05 - 1e-6).all()


def test_render_tile_matches_stage_pipeline(engine):
    """render_tile_all_gpu = synthesize(seed) -> expose(seed+1) -> develop -> scan(seed+2)
    (image_pipeline.rs:924-1014): the drop-in runs all four on the device (development
    and the developed-crystal BVH included); the stage API builds the scan's BVH on the
    host.  The device build reproduces the reference's partition order, so the two
    BVHs are the same node for node and leaf for leaf, and D_neg is the same bit for
    bit."""
    for seed, (w, h) in ((1234, (40, 32)), (77, (97, 61))):
        p = F.GpuTileParams.for_preset("tri-x", 2.4, w, h, exposure=1.0, tile_seed=seed)
        img = np.linspace(0.0, 1.0, w * h, dtype=np.float32)
        D = engine.render_tile_all_gpu(img, p).reshape(h, w)
        st = engine.last_tile_stats
        cfg = F.EmulsionConfig.tri_x_400(p.c.film_w, p.c.film_h)
        em = F.VirtualEmulsionLayer.synthesize(cfg, seed)
        F.expose_emulsion(em, img.reshape(h, w), F.ExposureConfig(6.0, 0.12, 1.3, 48), seed + 1, engine)
        rep = F.develop_emulsion(em, F.DevelopmentConfig(4, 1.2, 1.25), engine)
        D2, _ = F.scan_emulsion(em, F.ScanConfig(w, h, 4, F.ScanMode.NegativeDensity, 0.25), seed + 2, engine)
        assert st["crystals"] == len(em) and st["developed"] == rep.developed_crystals
        assert st["clumps"] == rep.total_clumps
        assert 0 < rep.developed_crystals < len(em)
        assert (D.view(np.uint32) == D2.view(np.uint32)).all()


def test_render_tile_all_gpu_from_threads_equals_sequential(engine):
    """Callers on several threads (the reference's Rayon tile loop, image_pipeline.rs:
    427-591) get exactly the results of sequential calls: the engine lock covers the GPU
    phases while synthesis runs on each caller's thread."""
    from concurrent.futures import ThreadPoolExecutor
    jobs = []
    for k in range(8):
        p = F.GpuTileParams.for_preset("tri-x", 2.4, 48, 40, exposure=1.0, tile_seed=500 + 17 * k)
        img = np.random.default_rng(k).uniform(0.0, 1.0, 48 * 40).astype(np.float32)
        jobs.append((img, p))
    seq = [engine.render_tile_all_gpu(img, p) for img, p in jobs]
    with ThreadPoolExecutor(4) as pool:
        par = list(pool.map(lambda j: engine.render_tile_all_gpu(*j), jobs))
    for a, b in zip(seq, par):
        assert (a.view(np.uint32) == b.view(np.uint32)).all()
