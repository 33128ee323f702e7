"""Benchmark of the photon-transport + development + scan forward on one MI355X.

Workload: frame24mp_trix_cubic -- a 6000x4000 frame as 247 tiles of 320 px (edge tiles
partial), Tri-X cubic emulsion per tile, 128 photons per pixel (3.072e9 photon
paths), supersample 2 (16 scan rays per pixel, 24 Mpix out).  One step renders the
whole frame: transport -> development -> scan -> OTF, with the per-tile emulsions
(crystals + BVH) already resident in HBM.  Prints one JSON line.

The line also carries the adjoint leg (tile512_trix_adjoint: expected-model forward and
its path-replay VJP, with their time ratio) and the CPU baseline.  The inverse-fit
workload prints its own line:

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode plain|batched]
  python bench.py --workload fit1024_trix1600_dye3 [--steps K]
  python bench.py --workload tile512_trix_adjoint [--steps K]
  python bench.py --workload film24mp_trix [--steps K]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_HBM_GBPS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sample_tiles(ranges, nx, ny, n=24):
    """A stratified sample of n tiles: four edge tiles (partial in this frame: top of the
    right column, middle of the right column, middle of the bottom row, the corner), the
    rest evenly spaced over the other tiles in raster order."""
    edge = sorted({nx - 1, (ny // 2) * nx + nx - 1, (ny - 1) * nx + nx // 2, ny * nx - 1})
    rest = [i for i in range(len(ranges)) if i not in edge]
    picks = [rest[int(round(k))] for k in np.linspace(0, len(rest) - 1, n - len(edge))]
    return sorted(set(edge + picks))


def cpu_baseline(cfg, exposure, frame, n_tiles=24, ppp_sample=32):
    """The oracle (C restatement of the reference's Rayon path: chunks = 8 * threads,
    row-parallel scan) timed on this host's cores over a stratified sample of the frame's
    tiles (edge tiles included), each at ppp_sample of the workload's photons per pixel:
    transport, development and scan timed per tile, synthesis excluded.  Projected to the
    frame by packets: transport time x (ppp / ppp_sample) + development + scan, summed over
    the sample and scaled by tiles / sampled tiles."""
    from oracle import oracle as O

    e, o, x = cfg["emulsion"], cfg["emulsion"]["optics"], cfg["exposure"]
    d, sc = cfg["development"], cfg["scan"]
    threads = os.cpu_count() or 1
    ranges = frame.tile_ranges()
    nx, ny = -(-frame.width // frame.tile), -(-frame.height // frame.tile)
    picks = sample_tiles(ranges, nx, ny, n_tiles)
    ppp = x["photons_per_pixel"]
    t_tr = t_dev = t_scan = 0.0
    paths = 0
    n_edge = 0
    for k in picks:
        x0, y0, tw, th, seed = ranges[k]
        n_edge += int(tw < frame.tile or th < frame.tile)
        cs = O.synthesize(W=tw, H=th, D=e["depth"], r_min=e["r_min"], mean_radius=e["mean_crystal_radius"],
                          size_sigma=e["size_variance"], morph=e["morphology"], aspect=e["tabular_aspect_ratio"],
                          base_sens=e["base_sensitivity"], sens_sigma=e["sensitivity_sigma"],
                          min_radius=e["min_radius"], min_half_thickness=e["min_half_thickness"], seed=seed)
        emap = np.ascontiguousarray(exposure[y0:y0 + th, x0:x0 + tw])
        t0 = time.perf_counter()
        counts, st = O.expose(cs, emap, W=tw, H=th, D=e["depth"], mu_a=o["absorption_coeff"],
                              mu_s=o["scattering_coeff"], g=o["anisotropy_g"], n_refr=o["refractive_index"],
                              r_ahu=o["antihalation_reflectance"], od_ahu=o["antihalation_density"],
                              mu_abs=x["crystal_absorption_coeff"], theta_max=x["illumination_angle_rad"],
                              max_bounces=x["max_bounces"], seed=seed + 1, sampler=0,
                              photons_per_area=float(ppp_sample), threads=threads,
                              base_thickness=x.get("base_thickness", 0.0))
        t1 = time.perf_counter()
        dev, fe, _, _ = O.develop(cs, counts, d["threshold_photons"], d["filament_expansion"],
                                  d["clumping_proximity_factor"])
        cs.developed, cs.fe = dev, fe
        t2 = time.perf_counter()
        O.scan(cs, tw, th, e["depth"], tw, th, spp=sc["samples_per_pixel"], ss=sc["supersample"],
               sigma=sc["silver_opacity"], seed=seed + 2, threads=threads)
        t3 = time.perf_counter()
        t_tr += t1 - t0
        t_dev += t2 - t1
        t_scan += t3 - t2
        paths += st["launched"]
    scale_p = ppp / ppp_sample
    frame_s = (t_tr * scale_p + t_dev + t_scan) * len(ranges) / len(picks)
    frame_paths = sum(tw * th for (_, _, tw, th, _) in ranges) * ppp
    return {
        "value": frame_paths / frame_s / 1e6,
        "unit": "Mphoton-paths/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "sample": (f"{len(picks)} of {len(ranges)} tiles ({n_edge} edge tiles), stratified over the frame, "
                   f"at {ppp_sample} of {ppp} photons per pixel ({paths} paths, reference chunking); "
                   f"transport x {scale_p:g} + development + scan per tile, x {len(ranges) / len(picks):.2f} "
                   f"to the frame; synthesis excluded; development and scan on the sample's smaller "
                   f"developed set"),
        "tiles": picks,
        "sample_s": {"transport": t_tr, "develop": t_dev, "scan": t_scan},
        "frame_s_projected": frame_s,
        "transport_mpaths_per_s": paths / t_tr / 1e6,
    }


def adjoint_leg(engine, name="tile512_trix_adjoint", reps=3):
    """Expected-model forward and its VJP (path replay) on the differentiation tile,
    beside the plain stochastic forward of the same tile (fg_expose -> fg_develop ->
    fg_scan with the same seeds)."""
    import numpy as np
    import fgrain_amd as F
    from fgrain_amd import workloads as W
    from fgrain_amd.adjoint import ExpectedFilm, mse_loss

    cfg = W.load_config(name)
    w, h, seed = cfg["frame"]["width"], cfg["frame"]["height"], cfg["seed"]
    em = F.VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, float(w), float(h)), seed)
    E = W.workload_exposure(cfg) * np.float32(W.packet_weight(cfg))
    film = ExpectedFilm(em, W.expected_model(cfg, w, h, seed), engine)
    tgt, _ = film.forward(E * np.float32(2.0 ** cfg["loss"]["target_shift_stops"]))
    out, dose = film.forward(E)
    L, G = mse_loss(out, tgt)
    film.vjp(E, G, dose=dose)
    tf, tb, ts = [], [], []
    kf, kb = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        out, dose = film.forward(E)
        t1 = time.perf_counter()
        kf.append(dict(film.last_stats))
        L, G = mse_loss(out, tgt)
        t2 = time.perf_counter()
        gE, pg = film.vjp(E, G, dose=dose)
        t3 = time.perf_counter()
        kb.append(dict(film.last_stats))
        tf.append((t1 - t0) * 1e3)
        tb.append((t3 - t2) * 1e3)
    zero = np.zeros(len(em), np.uint32)
    for _ in range(reps):
        em.set_state(counts=zero)
        t0 = time.perf_counter()
        F.expose_emulsion(em, E, W.exposure_config(cfg), seed + 1, engine)
        F.develop_emulsion(em, W.development_config(cfg), engine)
        F.scan_emulsion(em, W.scan_config(cfg, w, h), seed + 2, engine)
        ts.append((time.perf_counter() - t0) * 1e3)
    packets = kf[0]["packets"]
    mf, mb, ms = float(np.median(tf)), float(np.median(tb)), float(np.median(ts))
    return {
        "workload": name, "crystals": len(em), "packets": packets,
        "ms_forward": mf, "ms_vjp": mb, "adjoint_forward_ratio": mb / mf,
        "ms_plain_forward": ms, "adjoint_plain_forward_ratio": mb / ms,
        "ms_kernel_forward": {"replay": float(np.median([k["ms_replay"] for k in kf])),
                              "scan": float(np.median([k["ms_scan"] for k in kf]))},
        "ms_kernel_vjp": {"replay": float(np.median([k["ms_replay"] for k in kb])),
                          "scan_adjoint": float(np.median([k["ms_scan"] for k in kb]))},
        "replay_mpaths_per_s": packets / (float(np.median([k["ms_replay"] for k in kb])) * 1e-3) / 1e6,
        "loss": L, "grads": {"crystal_absorption_coeff": pg.crystal_absorption_coeff,
                             "anisotropy_g": pg.anisotropy_g,
                             "antihalation_reflectance": pg.antihalation_reflectance,
                             "exposure_l2": float(np.linalg.norm(gE))},
    }


def adjoint_main(args, engine):
    """Differentiation workload alone: one JSON line with the VJP / forward time ratio."""
    r = adjoint_leg(engine, args.workload, reps=args.steps)
    out = {
        "metric": "adjoint/forward time ratio (expected-model VJP by path replay over its forward)",
        "value": r["adjoint_forward_ratio"], "unit": "ratio", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": r["ms_forward"] + r["ms_vjp"], "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "packets": r["packets"], "crystals": r["crystals"]},
    }
    out.update({k: v for k, v in r.items() if k not in ("workload", "packets", "crystals")})
    print(json.dumps(out), flush=True)


def fit_main(args, engine, cfg):
    """Inverse fit workload: one JSON line with time per step, the adjoint/forward
    ratio and the plain-forward density RMSE."""
    from fgrain_amd.fit import DyeFit

    t = time.perf_counter()
    fit = DyeFit(engine, cfg, log=log)
    log(f"fit ready in {time.perf_counter() - t:.1f} s (host synthesis)")
    steps = args.steps if args.steps_given else cfg["fit"]["steps"]
    r = fit.run(steps=steps, log=log)
    fr = cfg["frame"]
    packets = fr["width"] * fr["height"] * cfg["exposure"]["photons_per_pixel"] * len(fit.layers)
    out = {
        "metric": "inverse fit: time per Adam step (expected forward + adjoint over all tiles)",
        "value": r["ms_per_step"], "unit": "ms/step", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": r["ms_per_step"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "frame": [fr["width"], fr["height"]],
                   "tile": fr["tile"], "layers": len(fit.layers),
                   "photons_per_pixel": cfg["exposure"]["photons_per_pixel"],
                   "packets_per_forward": packets},
        "adjoint_forward_ratio": r["adjoint_forward_ratio"],
        "ms_forward_per_step": r["ms_forward_per_step"], "ms_adjoint_per_step": r["ms_adjoint_per_step"],
        "loss_first": r["loss_first"], "loss_last": r["loss_last"],
        "rmse_plain_fit": r["rmse_plain_fit"], "rmse_plain_init": r["rmse_plain_init"],
        "rmse_plain_truth_resampled": r["rmse_plain_truth_resampled"],
        "exposure_rmse_log2": r["rel_exposure_rmse_log2"],
        "exposure_rmse_log2_init": r["rel_exposure_rmse_log2_init"],
    }
    print(json.dumps(out), flush=True)


def film_scene(cfg) -> "np.ndarray":
    """Seeded 8-bit RGB scene of the film workload: the blob luminance field with a warm
    tint, a black letterbox (top and bottom 8 %) and a black silhouette."""
    import numpy as np
    from fgrain_amd import workloads as W

    fr, im = cfg["film"], cfg["image"]
    w, h = fr["width"], fr["height"]
    lum = W.blob_exposure(w, h, cfg["seed"], blobs=im["blobs"], sigma=tuple(im["sigma"]),
                          amp=tuple(im["amplitude"]), lo=0.0, hi=1.0)
    rgb = np.empty((h, w, 3), np.uint8)
    for c, k in enumerate((1.0, 0.92, 0.8)):
        rgb[:, :, c] = np.clip(np.round(lum * 255.0 * k), 0, 255).astype(np.uint8)
    b = h * 8 // 100
    rgb[:b] = 0
    rgb[h - b:] = 0
    rgb[h // 2 - h // 20:h // 2 + h // 20, w // 3:w // 3 + w // 30] = 0
    return rgb


def film_main(args, engine, cfg):
    """Film render stages that are on (luma, gate + optics) at the workload's size: one
    JSON line with kernel times (HIP events), an HBM roofline per kernel and the oracle's
    single-thread CPU time for the same stages."""
    import numpy as np
    from fgrain_amd import film as FM

    rgb = film_scene(cfg)
    h, w = rgb.shape[:2]
    npx = w * h
    preset = cfg["film"]["preset"]
    t_luma, t_opt = [], []
    gray = FM.film_luma(rgb, preset, engine)          # warm-up
    FM.film_optics(gray, engine)
    for _ in range(args.steps):
        tm = {}
        gray = FM.film_luma(rgb, preset, engine, timing=tm)
        t_luma.append(tm["ms_kernel"])
        heff = FM.film_optics(gray, engine, timing=tm)
        t_opt.append(tm["ms_kernel"])
        log(f"luma {t_luma[-1]:.3f} ms, gate + optics {t_opt[-1]:.3f} ms")
    ms_l, ms_o = float(np.median(t_luma)), float(np.median(t_opt))
    if not (ms_l > 0 and ms_o > 0):
        raise SystemExit(f"film stages: no kernel time from the HIP events (luma {ms_l}, optics {ms_o})")
    # algorithmic bytes: luma reads RGB and writes gray; the gate reads gray, the optics
    # reads gray (3-pixel halo from LDS) and writes h_effective f32
    b_l, b_o = 4.0 * npx, 6.0 * npx
    gb_l, gb_o = b_l / ms_l / 1e6, b_o / ms_o / 1e6
    out = {
        "metric": "film render stages on the GPU: Mpix/s of luma + gate + optics (24 MP)",
        "value": npx / (ms_l + ms_o) / 1e3, "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": ms_l + ms_o, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8/f32", "data": "synthetic",
        "config": {"workload": args.workload, "frame": [w, h], "preset": preset},
        "ms_kernel": {"luma": ms_l, "gate_optics": ms_o},
        "roofline": {"bound": "hbm", "achieved": gb_o, "peak": MI355X_HBM_GBPS, "unit": "GB/s",
                     "frac": gb_o / MI355X_HBM_GBPS, "traffic": None, "kernel": "gate + optics",
                     "bytes_per_px": 6.0},
        "roofline_luma": {"bound": "hbm", "achieved": gb_l, "peak": MI355X_HBM_GBPS, "unit": "GB/s",
                          "frac": gb_l / MI355X_HBM_GBPS, "bytes_per_px": 4.0},
        "compositing": "switched off in this build (device memory fault under investigation)",
        "h_effective_mean": float(heff.mean()),
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        t = time.perf_counter()
        g2 = O.film_luma(rgb, preset)
        t1 = time.perf_counter()
        O.film_optics(g2)
        t2 = time.perf_counter()
        out["cpu_baseline"] = {"value": npx / (t2 - t) / 1e6, "unit": "Mpix/s", "cores": 1, "kind": "port",
                               "sample": f"the whole {w}x{h} frame, luma {t1 - t:.2f} s + optics {t2 - t1:.2f} s"}
    print(json.dumps(out), flush=True)


def valu_probe(iters, blocks=4096):
    """The VALU calibration kernel alone (blocks x 256 threads, 64 * iters + 4 VALU
    instructions per wave): one JSON line with its known instruction count and time, for
    scaling rocprofv3's SQ_INSTS_VALU and reading the device's VALU issue rate."""
    import ctypes as C
    import re

    import fgrain_amd as F
    from fgrain_amd import _lib as L

    engine = F.GpuEngine.new()
    m = re.search(r"(\d+) CUs", engine.adapter_name())
    cus = int(m.group(1)) if m else 256
    ms, per_wave = C.c_float(), C.c_uint64()
    for _ in range(3):  # the last of three runs
        L.check(L.lib().fg_calibrate_valu(engine._h, blocks, iters, C.byref(ms), C.byref(per_wave)))
    waves = blocks * 4
    insts = waves * per_wave.value
    rate = insts / (ms.value * 1e-3)                 # wave instructions per second
    print(json.dumps({"valu_probe_blocks": blocks, "waves": waves, "valu_per_wave": per_wave.value,
                      "valu_insts": insts, "ms_kernel": ms.value, "cus": cus,
                      "insts_per_s": rate,
                      "insts_per_simd_per_ns": rate / (cus * 4) / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["plain", "batched"], default="batched")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-tile comparison step")
    ap.add_argument("--workload", default="frame24mp_trix_cubic")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--synthesis", choices=["host", "device"], default=None,
                    help="emulsion synthesis of the frame (default: the workload's, host Bridson)")
    ap.add_argument("--no-tuning", action="store_true",
                    help="ignore fgrain_amd/configs/tuning/<workload>.json (autotuner output)")
    ap.add_argument("--set-tuning", action="append", default=[], metavar="KEY=VALUE",
                    help="override one tuning field after the workload's tuning is loaded")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 4 times (counter calibration)")
    ap.add_argument("--valu-probe", type=int, default=0, metavar="ITERS",
                    help="only run the VALU calibration kernel (64 * ITERS + 4 VALU per wave)")
    args = ap.parse_args()
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 3
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a frame's tiles would shard across GPUs, a contiguous range of tiles per rank)",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        a = torch.empty(args.copy_probe << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe << 20, "copies": 4}))
        return

    import fgrain_amd as F
    from fgrain_amd import workloads as W

    from fgrain_amd.autotune import load_tuning

    if args.valu_probe:
        return valu_probe(args.valu_probe)

    cfg = W.load_config(args.workload)
    engine = F.GpuEngine.new()
    log(f"engine: {engine.adapter_name()}")
    tuned = None if args.no_tuning else load_tuning(args.workload)
    if tuned:
        engine.set_tuning(**tuned["tuning"])
        log(f"tuning ({tuned['measured']}): {tuned['tuning']}")
    if args.set_tuning:
        t = engine.tuning()
        for kv in args.set_tuning:
            k, _, v = kv.partition("=")
            if k not in t or not v:
                raise SystemExit(f"bench.py: --set-tuning {kv!r}: expected one of {sorted(t)}=<int>")
            t[k] = int(v)
        engine.set_tuning(**t)
        log(f"tuning overridden: {t}")
    if "fit" in cfg:
        return fit_main(args, engine, cfg)
    if "film" in cfg:
        return film_main(args, engine, cfg)
    if "loss" in cfg:
        return adjoint_main(args, engine)
    t = time.perf_counter()
    exposure = W.workload_exposure(cfg)
    frame = W.Frame(engine, cfg, exposure, synthesis=args.synthesis)
    t_synth = time.perf_counter() - t
    log(f"frame ready: {frame.crystals} crystals in {t_synth:.1f} s ({frame.synthesis} synthesis)")
    mode = W.Frame.BATCHED if args.mode == "batched" else W.Frame.PLAIN
    for i in range(args.warmup):
        st = frame.render(mode)
        log(f"warmup {i}: {st['ms_total']:.1f} ms")
    agg = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = frame.render(mode)
        log(f"step {i}: {st['ms_total']:.1f} ms (transport kernels {st['ms_kernel_transport']:.1f} ms, "
            f"scan kernels {st['ms_kernel_scan']:.1f} ms)")
        if agg is None:
            agg = dict(st)
        else:
            for k in ("ms_kernel_transport", "ms_kernel_scan", "ms_transport", "ms_develop", "ms_scan"):
                agg[k] += st[k]
            agg["graph_used"] = min(agg["graph_used"], st["graph_used"])
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / args.steps * 1e3
    photons = agg["photons"]
    pixels = agg["pixels"]
    # Dominant kernel: the frame transport, one launch per step.  Its HBM roofline is taken
    # on the bytes it actually moves to and from HBM: rocprofv3 FETCH_SIZE + WRITE_SIZE of
    # that kernel in one frame, each scaled by what the counter reads for a copy of known
    # size (profiles/traffic_<workload>.json), over the kernel's live HIP-event time, so
    # the fraction cannot exceed 1.  The reference algorithm's bytes per path (oracle count
    # over sampled tiles: 4 B exposure + 32 B per BVH node + 48 B per crystal record + 8 B
    # per absorbing hit on the full infinite-ray traversal) are reported beside it: they
    # are served from L2 and the MALL, not HBM.  What bounds the kernel is instruction
    # issue: the VALU issue fraction comes from the same committed profile (SQ_INSTS_VALU
    # checked against the VALU calibration kernel of known instruction count, which also
    # measures the issue rate: about 2 SIMD cycles per wave64 f32 instruction on gfx950).
    per_path = json.load(open(os.path.join(ROOT, "tests", "golden", "roofline_bytes.json"))).get(
        args.workload, {}).get("transport_per_path")
    if per_path is None:
        raise SystemExit(f"no algorithmic byte count for {args.workload}: "
                         f"run tests/golden/make_roofline_bytes.py {args.workload}")
    alg_bytes = per_path["bytes"] * photons
    dev_bytes = 4 * photons + 32 * agg["nodes_visited"] + 48 * agg["prims_tested"] + 8 * agg["abs_events"]
    prof = {}
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath) and args.mode == "batched":
        prof = json.load(open(tpath))
    ms_kernel = agg["ms_kernel_transport"] / args.steps
    roof = None
    if ms_kernel > 0:
        alg_gbps = alg_bytes / (ms_kernel * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": None, "peak": MI355X_HBM_GBPS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": prof.get("kernel", "transport"),
                "launches_per_step": 1 if args.mode == "batched" else agg["tiles"],
                "algorithmic": {"bytes_per_path": per_path["bytes"], "bytes": alg_bytes,
                                "gbps": alg_gbps, "frac_of_hbm": alg_gbps / MI355X_HBM_GBPS,
                                "note": "reference traversal bytes, served from L2/MALL"}}
        if prof:
            traffic = prof["transport_traffic_bytes"]
            ach = traffic / (ms_kernel * 1e-3) / 1e9
            roof.update({"achieved": ach, "frac": ach / MI355X_HBM_GBPS, "traffic": traffic,
                         "traffic_profile_ms": prof.get("transport_kernel_ms")})
            roof["issue"] = {k: prof.get(k) for k in (
                "transport_valu_insts_per_path", "transport_valu_issue_frac",
                "transport_valu_issue_frac_of_probe", "transport_valu_lane_util", "transport_l2_hit_rate",
                "valu_counter_scale")}
    out = {
        "metric": "Mphoton-paths/s + scanned Mpix/s, 24MP Tri-X, 1 GPU; adjoint/forward time ratio",
        "value": photons / (ms_step * 1e-3) / 1e6,
        "unit": "Mphoton-paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "frame": [frame.width, frame.height],
                   "tiles": agg["tiles"], "photons_per_pixel": cfg["exposure"]["photons_per_pixel"],
                   "supersample": cfg["scan"]["supersample"], "mode": args.mode,
                   "crystals": agg["crystals"], "synthesis": frame.synthesis},
        "s_frame_synthesis": t_synth,
        "scanned_mpix_per_s": pixels / (ms_step * 1e-3) / 1e6,
        "transport_kernel_mpaths_per_s": photons * args.steps / (agg["ms_kernel_transport"] * 1e-3) / 1e6,
        "scan_kernel_mpix_per_s": pixels * args.steps / (agg["ms_kernel_scan"] * 1e-3) / 1e6,
        "ms_breakdown": {k: agg[k] / args.steps for k in ("ms_transport", "ms_develop", "ms_scan",
                                                          "ms_kernel_transport", "ms_kernel_scan")},
        "roofline": roof,
        "graph_used": agg.get("graph_used"),
        "transport_atomics": agg.get("atomics"),
        "transport_per_path_executed": {
            "bytes": dev_bytes / photons, "nodes": agg["nodes_visited"] / photons,
            "prims": agg["prims_tested"] / photons, "bounces": agg["bounces"] / photons},
        "transport_per_path_reference": per_path,
        "tuning": engine.tuning(),
    }
    if args.mode == "batched" and not args.no_plain:
        # the plain per-tile path on the same resident frame: per tile one launch sequence
        # (transport, device development + developed-crystal BVH, scan) and a
        # synchronisation; its one-tile frames are built by an untimed first render
        log("plain per-tile comparison: set-up render ...")
        frame.render(W.Frame.PLAIN)
        log("plain per-tile comparison step ...")
        t1 = time.perf_counter()
        sp = frame.render(W.Frame.PLAIN)
        plain_ms = (time.perf_counter() - t1) * 1e3
        out["plain_ms_per_step"] = plain_ms
        out["batched_speedup_vs_plain"] = plain_ms / ms_step
        out["plain_kernels_ms"] = {"transport": sp["ms_kernel_transport"], "develop": sp["ms_develop"],
                                   "scan": sp["ms_kernel_scan"]}
        out["batched_kernels_ms"] = {k: agg[m] / args.steps for k, m in (
            ("transport", "ms_kernel_transport"), ("develop", "ms_develop"), ("scan", "ms_kernel_scan"))}
        out["batched_speedup_vs_plain_kernels"] = (
            sum(out["plain_kernels_ms"].values()) / sum(out["batched_kernels_ms"].values()))
    if not args.no_adjoint:
        log("adjoint leg (tile512_trix_adjoint) ...")
        out["adjoint"] = adjoint_leg(engine)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, host cores) ...")
        out["cpu_baseline"] = cpu_baseline(cfg, exposure, frame)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
