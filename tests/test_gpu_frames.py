"""Frames on the GPU: the plain per-tile path (one launch sequence and synchronisation
per tile: transport, device development and developed-crystal BVH, scan) equals the
stage pipeline tile by tile, the batched device path (one persistent transport over all
tiles, the same development and scan kernels over all tiles, hipGraphs) equals the plain
path bit for bit, graph replays are deterministic, and device development equals the
oracle."""
import copy

import numpy as np
import pytest

import fgrain_amd as F
from fgrain_amd import workloads as W
from fgrain_amd._lib import GpuError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cfg(ppp=16, otf=0, ss=2, seed=5):
    cfg = copy.deepcopy(W.load_config("frame24mp_trix_cubic"))
    cfg["exposure"]["photons_per_pixel"] = ppp
    cfg["scan"]["otf"] = otf
    cfg["scan"]["supersample"] = ss
    cfg["seed"] = seed
    return cfg


@pytest.fixture(scope="module")
def frame(engine):
    cfg = _cfg()
    w, h = 680, 420  # 3 x 2 tiles of 320 px, partial edge tiles 40 and 100 px
    exp = W.blob_exposure(w, h, 5, sigma=(40, 160))
    return cfg, W.Frame(engine, cfg, exp, frame=(w, h))


def _tile_stage(engine, cfg, exp, x0, y0, tw, th, seed):
    em = F.VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, float(tw), float(th)), seed)
    F.expose_emulsion(em, np.ascontiguousarray(exp[y0:y0 + th, x0:x0 + tw]), W.exposure_config(cfg),
                      seed + 1, engine)
    F.develop_emulsion(em, W.development_config(cfg), engine)
    D, _ = F.scan_emulsion(em, W.scan_config(cfg, tw, th), seed + 2, engine)
    return em, D


def test_plain_frame_equals_stage_pipeline_per_tile(engine, frame):
    """Per tile, the frame's plain path (device development and developed-crystal BVH)
    and the stage API (the scan's BVH built on the host) deposit the same photons,
    develop the same crystals and give the same D_neg bit for bit: the device build
    reproduces the reference's partition order."""
    cfg, fr = frame
    st = fr.render(W.Frame.PLAIN)
    assert st["tiles"] == 6 and st["photons"] == 16 * 680 * 420
    D = fr.read()
    counts = fr.counts()
    dev, fe, clump = fr.development()
    base = 0
    for (x0, y0, tw, th, seed) in fr.tile_ranges():
        em, Dt = _tile_stage(engine, cfg, fr.exposure, x0, y0, tw, th, seed)
        n = len(em)
        s = em.state()
        assert np.array_equal(counts[base:base + n], em.absorbed_photons())
        assert np.array_equal(dev[base:base + n], s["developed"]) and np.array_equal(fe[base:base + n], s["fe"])
        assert np.array_equal(clump[base:base + n], s["clump_id"])
        assert (D[y0:y0 + th, x0:x0 + tw].view(np.uint32) == Dt.view(np.uint32)).all()
        base += n
    assert base == fr.crystals


def test_batched_frame_equals_plain(engine, frame):
    cfg, fr = frame
    sp = fr.render(W.Frame.PLAIN)
    Dp, cp, (dp, fp, kp) = fr.read(), fr.counts(), fr.development()
    sb = fr.render(W.Frame.BATCHED)
    Db, cb, (db, fb, kb) = fr.read(), fr.counts(), fr.development()
    assert np.array_equal(cp, cb)
    assert np.array_equal(dp, db)
    assert np.array_equal(kp, kb)
    assert np.array_equal(fp, fb)
    assert sp["developed"] == sb["developed"] and sp["clumps"] == sb["clumps"]
    assert sp["nodes_visited"] == sb["nodes_visited"] and sp["prims_tested"] == sb["prims_tested"]
    # the same development and scan kernels per tile: the same image bit for bit
    assert (Dp.view(np.uint32) == Db.view(np.uint32)).all()


def test_batched_graph_replay_is_deterministic(engine, frame):
    cfg, fr = frame
    outs = []
    for _ in range(3):
        st = fr.render(W.Frame.BATCHED)
        assert st["graph_used"] == 1      # the replayed hipGraph, not direct launches
        outs.append((fr.read(), fr.counts()))
    assert fr.render(W.Frame.PLAIN)["graph_used"] == 0
    for D, c in outs[1:]:
        assert (D.view(np.uint32) == outs[0][0].view(np.uint32)).all()
        assert np.array_equal(c, outs[0][1])


def test_batched_transport_occupancy_gives_identical_frames(engine, frame):
    """The 7-waves-per-SIMD build of the frame transport (fewer VGPRs, more spills) and
    the 6-wave build deposit the same photons and give the same frame bit for bit."""
    cfg, fr = frame
    old = engine.tuning()
    outs = []
    try:
        for waves in (6, 7):
            engine.set_tuning(**dict(old, transport_waves=waves, transport_wait_pct=90))
            assert engine.tuning()["transport_waves"] == waves
            st = fr.render(W.Frame.BATCHED)
            outs.append((fr.read(), fr.counts(), st["nodes_visited"], st["prims_tested"]))
        with pytest.raises(GpuError, match="transport_waves"):
            engine.set_tuning(**dict(old, transport_waves=5))
    finally:
        engine.set_tuning(**old)
    (D6, c6, n6, p6), (D7, c7, n7, p7) = outs
    assert np.array_equal(c6, c7) and n6 == n7 and p6 == p7
    assert (D6.view(np.uint32) == D7.view(np.uint32)).all()


def test_deposit_prereduction_gives_identical_counts_with_fewer_atomics(engine, frame):
    """The wave pre-reduction of deposits (record_absorption, crystal.rs:221-223, as one
    atomic per group of lanes depositing into the same crystal) gives the same latent
    counts and frame bit for bit, in both occupancy builds of the frame transport and in
    the per-tile transport, with fewer global atomics than absorbing hits."""
    cfg, fr = frame
    old = engine.tuning()
    outs = {}
    try:
        for waves in (6, 7):
            for pre in (0, 1):
                engine.set_tuning(**dict(old, transport_waves=waves, transport_prereduce=pre))
                st = fr.render(W.Frame.BATCHED)
                outs[waves, pre] = (fr.read(), fr.counts(), st["abs_events"], st["atomics"])
        with pytest.raises(GpuError, match="transport_prereduce"):
            engine.set_tuning(**dict(old, transport_prereduce=2))
        x0, y0, tw, th, seed = fr.tile_ranges()[0]
        tile_counts = []
        for pre in (0, 1):
            engine.set_tuning(**dict(old, transport_prereduce=pre))
            em = F.VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, float(tw), float(th)), seed)
            st = F.expose_emulsion(em, np.ascontiguousarray(fr.exposure[y0:y0 + th, x0:x0 + tw]),
                                   W.exposure_config(cfg), seed + 1, engine)
            tile_counts.append((em.absorbed_photons(), st["abs_events"], st["atomics"]))
    finally:
        engine.set_tuning(**old)
    D0, c0, ev0, at0 = outs[6, 0]
    assert at0 == ev0 > 0
    for key, (D, c, ev, at) in outs.items():
        assert np.array_equal(c, c0), key
        assert (D.view(np.uint32) == D0.view(np.uint32)).all(), key
        assert ev == ev0
        if key[1] == 1:
            assert 0 < at < ev, key
    (a0, e0, t0), (a1, e1, t1) = tile_counts
    assert np.array_equal(a0, a1) and e0 == e1 and t0 == e0 and 0 < t1 < e1


def test_batched_development_matches_oracle(engine):
    cfg = _cfg(ppp=2, seed=11)
    w, h = 400, 200
    exp = W.ramp_exposure(w, h, 0.01, 1.5)
    fr = W.Frame(engine, cfg, exp, frame=(w, h))
    fr.render(W.Frame.BATCHED)
    counts = fr.counts()
    dev, fe, clump = fr.development()
    d = cfg["development"]
    base = 0
    for (x0, y0, tw, th, seed) in fr.tile_ranges():
        em = F.VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, float(tw), float(th)), seed)
        n = len(em)
        c = em.crystals()
        cs = O.Crystals(c["pos"], c["dims"], c["quat"], c["morph"], c["sens"])
        od, ofe, ocid, rep = O.develop(cs, counts[base:base + n], d["threshold_photons"],
                                       d["filament_expansion"], d["clumping_proximity_factor"])
        assert 0 < rep["developed"] < n and rep["clumps"] > 1
        assert np.array_equal(dev[base:base + n], od)
        assert np.array_equal(clump[base:base + n], ocid)
        assert np.array_equal(fe[base:base + n], ofe)
        base += n


@pytest.mark.parametrize("threshold", [0, 1 << 30])
def test_batched_development_extremes_equal_plain(engine, threshold):
    """Every crystal developed (the wave-aggregated clump sizes and per-tile largest clump,
    every BVH size class and the bottom-up leaf counts at full load) and none developed
    (empty development lists, T = 1): the batched device path equals the per-tile path."""
    cfg = _cfg(ppp=2, seed=21)
    cfg["development"]["threshold_photons"] = threshold
    w, h = 360, 200
    fr = W.Frame(engine, cfg, W.blob_exposure(w, h, 3, sigma=(30, 90)), frame=(w, h))
    sp = fr.render(W.Frame.PLAIN)
    Dp, (dp, fp, kp) = fr.read(), fr.development()
    sb = fr.render(W.Frame.BATCHED)
    Db, (db, fb, kb) = fr.read(), fr.development()
    assert np.array_equal(dp, db) and np.array_equal(kp, kb) and np.array_equal(fp, fb)
    assert sp["developed"] == sb["developed"] and sp["clumps"] == sb["clumps"]
    if threshold == 0:
        assert sb["developed"] == fr.crystals and 0 < sb["clumps"] <= sb["developed"]
    else:
        assert sb["developed"] == 0 and (Db == np.float32(0.05)).all()
    assert np.abs(Dp - Db).max() <= 1e-5


def test_batched_frame_with_otf_matches_oracle_filter(engine):
    cfg = _cfg(ppp=4, otf=1, ss=1, seed=3)
    w, h = 352, 90
    fr = W.Frame(engine, cfg, W.blob_exposure(w, h, 3, sigma=(20, 80)), frame=(w, h))
    fr.render(W.Frame.BATCHED)
    D_otf = fr.read()
    cfg2 = _cfg(ppp=4, otf=0, ss=1, seed=3)
    fr2 = W.Frame(engine, cfg2, fr.exposure, frame=(w, h))
    fr2.render(W.Frame.BATCHED)
    assert (D_otf.view(np.uint32) == O.otf3(fr2.read()).view(np.uint32)).all()


def test_tmax_frame_with_uneven_tiles_batched_equals_plain(engine):
    """T-Max tabular tiles with per-tile density multipliers (uneven crystal counts),
    halation below a film base: one batched device pass equals the per-tile path."""
    cfg = copy.deepcopy(W.load_config("frame24mp_tmax_tabular"))
    cfg["exposure"]["photons_per_pixel"] = 8
    w, h = 640, 330
    mult = np.array([0.5, 3.0, 1.2, 2.2, 0.8, 1.0], np.float32)
    fr = W.Frame(engine, cfg, W.blob_exposure(w, h, 2, sigma=(40, 160)), density_mult=mult, frame=(w, h))
    sp = fr.render(W.Frame.PLAIN)
    Dp, cp, (dp, fp, kp) = fr.read(), fr.counts(), fr.development()
    sb = fr.render(W.Frame.BATCHED)
    Db, cb, (db, fb, kb) = fr.read(), fr.counts(), fr.development()
    assert np.array_equal(cp, cb) and np.array_equal(dp, db) and np.array_equal(kp, kb)
    assert np.array_equal(fp, fb)
    assert np.abs(Dp - Db).max() <= 1e-5
    assert sp["developed"] == sb["developed"] > 0


def test_scan_shape_changes_only_the_ray_sum_order(engine, frame):
    cfg, fr = frame
    old = engine.tuning()
    try:
        outs = []
        for rpl in (0, 2, 4, 16):
            engine.set_tuning(**dict(old, scan_rays_per_thread=rpl))
            fr.render(W.Frame.BATCHED)
            outs.append((fr.read(), fr.counts()))
    finally:
        engine.set_tuning(**old)
    for D, c in outs[1:]:
        assert np.array_equal(c, outs[0][1])
        assert np.abs(D - outs[0][0]).max() <= 1e-5


def test_autotuner_on_a_small_crop(engine, monkeypatch):
    """The tuner times every candidate, checks that transport candidates reproduce the
    default latent counts, and returns the fastest of each kernel's knobs."""
    from fgrain_amd import autotune as A
    monkeypatch.setattr(A, "TRANSPORT_GRID", {"transport_chunk": [64, 256], "transport_wait_pct": [80, 100]})
    monkeypatch.setattr(A, "SCAN_RAYS_PER_LANE", [1, 4])
    r = A.tune(engine, "tile320_trix_cubic", tiles=(1, 1), reps=1, log=lambda m: None)
    assert len(r["results"]) == 4 + 2
    best = min(x["ms_kernel_transport"] for x in r["results"] if "ms_kernel_transport" in x)
    assert r["ms_kernel_transport"] == best
    assert r["tuning"]["scan_rays_per_thread"] in (1, 4)
    assert engine.tuning()["transport_chunk"] == 64   # restored


@pytest.mark.parametrize("workload", ["frame24mp_trix_cubic", "frame24mp_tmax_tabular"])
def test_full_frame_properties(engine, workload):
    """The benchmarks' full frames (6000 x 4000 px, 247 tiles, 3.07e9 photon paths; Tri-X
    cubic, and T-Max tabular with uneven tile densities and a film base) on the batched
    path, through properties that do not depend on their size: graph replays are
    bit-identical, another launch tuning (packet chunk, leaf wait, transport occupancy)
    deposits the same photons, every latent photon belongs to a counted absorption
    event, and D_neg stays inside its mapped range."""
    cfg = W.load_config(workload)
    fr = W.Frame(engine, cfg, W.workload_exposure(cfg))
    old = engine.tuning()
    try:
        engine.set_tuning(**dict(old, transport_chunk=256, transport_wait_pct=90, transport_waves=7))
        s1 = fr.render(W.Frame.BATCHED)
        D1, c1 = fr.read(), fr.counts()
        fr.render(W.Frame.BATCHED)
        D2, c2 = fr.read(), fr.counts()
        engine.set_tuning(**dict(old, transport_chunk=64, transport_wait_pct=80, transport_waves=6))
        s3 = fr.render(W.Frame.BATCHED)
        c3 = fr.counts()
    finally:
        engine.set_tuning(**old)
    assert s1["tiles"] == 247 and s1["photons"] == 128 * 6000 * 4000 and s1["rays"] == 16 * 6000 * 4000
    # every packet of the frame was started (packet ids run past 2^31 here)
    assert s1["packets_run"] == s1["photons"] and s3["packets_run"] == s3["photons"]
    assert np.array_equal(c1, c2) and (D1.view(np.uint32) == D2.view(np.uint32)).all()
    assert np.array_equal(c1, c3)
    assert (s3["bounces"], s3["nodes_visited"], s3["prims_tested"], s3["abs_events"]) == \
        (s1["bounces"], s1["nodes_visited"], s1["prims_tested"], s1["abs_events"])
    assert s1["developed"] == s3["developed"] and 0 < s1["developed"] < s1["crystals"]
    assert int(c1.sum(dtype=np.uint64)) >= s1["abs_events"] > 0
    assert np.isfinite(D1).all()
    # the scanner OTF is a convex 3-tap blend, so it keeps the range up to rounding
    assert D1.min() >= 0.05 - 1e-5 and D1.max() <= 4.05 + 1e-5


def test_packet_ids_past_2_pow_31_batched_equals_plain(engine):
    """A frame of 2.2e9 packets (two tiles at 10600 photons per pixel): the batched
    transport's frame-wide packet ids pass 2^31 (a sign-extended 32-bit id would drop
    every packet beyond it), and a tile of 2.2e9 packets takes the ids of the one-tile
    frames and of the stage API's transport (fg_expose) past 2^31 too.  Every packet is
    started, and the counts and images of the batched and the per-tile path are
    identical."""
    cfg = _cfg(ppp=10600, seed=41)
    for (w, h), ppp in (((640, 320), 10600), ((320, 320), 21000)):
        cfg["exposure"]["photons_per_pixel"] = ppp
        print(f"frame {w}x{h} ppp {ppp}: {w * h * ppp} packets", flush=True)
        fr = W.Frame(engine, cfg, W.ramp_exposure(w, h, 0.01, 0.05), frame=(w, h))
        sb = fr.render(W.Frame.BATCHED)
        cb, Db = fr.counts(), fr.read()
        sp = fr.render(W.Frame.PLAIN)
        cp, Dp = fr.counts(), fr.read()
        assert sb["photons"] > 2 ** 31
        assert sb["packets_run"] == sb["photons"] == sp["packets_run"]
        assert np.array_equal(cb, cp)
        assert (Db.view(np.uint32) == Dp.view(np.uint32)).all()
        if len(fr.tile_ranges()) == 1:
            x0, y0, tw, th, seed = fr.tile_ranges()[0]
            em = F.VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, float(tw), float(th)), seed)
            st = F.expose_emulsion(em, np.ascontiguousarray(fr.exposure), W.exposure_config(cfg), seed + 1, engine)
            assert st["packets_run"] == sb["photons"]
            assert np.array_equal(em.absorbed_photons(), cb)


@pytest.mark.parametrize("size", [(7, 5), (33, 2), (1, 1)])
def test_tiny_batched_frame_equals_plain(engine, size):
    """Frames of a few pixels (one partial tile, a handful of crystals, at most a few
    developed): the batched device path's counters, lists and developed-crystal BVH at
    their smallest sizes equal the per-tile path."""
    w, h = size
    cfg = _cfg(ppp=64, seed=29)
    fr = W.Frame(engine, cfg, np.full((h, w), 3.0, np.float32), frame=(w, h))
    sp = fr.render(W.Frame.PLAIN)
    Dp, cp, (dp, fp, kp) = fr.read(), fr.counts(), fr.development()
    sb = fr.render(W.Frame.BATCHED)
    Db, cb, (db, fb, kb) = fr.read(), fr.counts(), fr.development()
    assert sb["tiles"] == 1 and sb["photons"] == 64 * w * h
    assert np.array_equal(cp, cb) and np.array_equal(dp, db) and np.array_equal(kp, kb)
    assert np.array_equal(fp, fb)
    assert sp["developed"] == sb["developed"] and sp["clumps"] == sb["clumps"]
    assert np.abs(Dp - Db).max() <= 1e-5


def test_cubic_frame_with_uneven_tiles_batched_equals_plain(engine):
    """Tri-X cubic tiles at 3x to 0.5x the crystal density have different largest
    semi-axes (r_min * 0.48), so their corners protrude by different reaches beyond the
    reference's AABBs; the first tile is the densest (smallest reach).  The batched
    launch prunes every tile with the largest reach, the per-tile path each with its
    own: the same photons, development and image bit for bit."""
    cfg = _cfg(ppp=16, seed=61)
    w, h = 640, 330
    mult = np.array([3.0, 0.5, 1.2, 0.8, 2.0, 1.0], np.float32)
    fr = W.Frame(engine, cfg, W.blob_exposure(w, h, 3, sigma=(40, 120)), density_mult=mult, frame=(w, h))
    sp = fr.render(W.Frame.PLAIN)
    Dp, cp, (dp, fp, kp) = fr.read(), fr.counts(), fr.development()
    sb = fr.render(W.Frame.BATCHED)
    Db, cb, (db, fb, kb) = fr.read(), fr.counts(), fr.development()
    assert np.array_equal(cp, cb) and np.array_equal(dp, db) and np.array_equal(kp, kb)
    assert np.array_equal(fp, fb)
    assert (Dp.view(np.uint32) == Db.view(np.uint32)).all()
    assert sp["developed"] == sb["developed"] > 0
