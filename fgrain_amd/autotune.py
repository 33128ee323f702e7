"""Autotuner of the engine's launch parameters (fg_tuning) for a frame workload.

Knobs (include/fgrain_hip.h, fg_tuning) that the batched frame path uses (it always runs
the persistent lane-regenerating transport; the kernel choice applies to per-tile calls):
* photon-packet fetch size: packets a wave takes from the work counter at a time (chunk);
* leaf-batch threshold: BVH node steps continue until this percentage of the walking
  lanes hold a full leaf queue (wait_pct);
* transport occupancy: 6 or 7 waves per SIMD (register budget against latency hiding);
* deposit pre-reduction across the wave before the global atomics (on or off);
* scan tile shape: rays of a pixel per lane (the pixel's rays / that many adjacent lanes).
The BVH walk is stackless (threaded skip links), so there is no traversal-stack depth to
tune.  Transport and scan are separate kernels, so their knobs are searched separately:
each candidate renders a crop of the workload's frame (whole tiles, edge tiles included)
on the batched path and is scored by the median HIP-event time of its kernel.  Transport
candidates must reproduce the default's latent counts exactly (they only reorder work).

    python -m fgrain_amd.autotune --workload frame24mp_trix_cubic [--tiles 8x6] [--reps 3]

writes fgrain_amd/configs/tuning/<workload>.json, which bench.py applies.
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs", "tuning")

TRANSPORT_GRID = {
    "transport_chunk": [64, 128, 256, 512],
    "transport_wait_pct": [65, 80, 90, 95, 100],
    "transport_waves": [6, 7],
    "transport_prereduce": [0, 1],
}
SCAN_RAYS_PER_LANE = [1, 2, 4, 8, 16]


def tuning_path(workload: str) -> str:
    return os.path.join(TUNING_DIR, workload + ".json")


def load_tuning(workload: str) -> dict | None:
    p = tuning_path(workload)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def _median_ms(frame, key, reps):
    from .workloads import Frame
    frame.render(Frame.BATCHED)  # warm-up: rebuild for the new tuning, capture the graph
    return float(np.median([frame.render(Frame.BATCHED)[key] for _ in range(reps)]))


def tune(engine, workload: str, tiles=(8, 6), reps=3, log=print) -> dict:
    from . import workloads as W

    cfg = W.load_config(workload)
    t = cfg["frame"]["tile"]
    w = min(cfg["frame"]["width"], tiles[0] * t - t // 4)   # last column partial
    h = min(cfg["frame"]["height"], tiles[1] * t - t // 2)  # last row partial
    exposure = np.ascontiguousarray(W.workload_exposure(cfg)[:h, :w])
    frame = W.Frame(engine, cfg, exposure, frame=(w, h))
    log(f"autotune {workload}: crop {w}x{h} px, {len(frame.tile_ranges())} tiles, {frame.crystals} crystals")
    default = {"transport_kernel": 0, "transport_chunk": 64, "transport_wait_pct": 80, "scan_rays_per_thread": 0,
               "transport_waves": 6, "transport_prereduce": 0}
    engine.set_tuning(**default)
    frame.render(W.Frame.BATCHED)
    ref_counts = frame.counts().copy()
    results = []
    best_t, best_ms = None, float("inf")
    keys = list(TRANSPORT_GRID)
    for values in itertools.product(*(TRANSPORT_GRID[k] for k in keys)):
        cand = dict(default, **dict(zip(keys, values)))
        engine.set_tuning(**cand)
        ms = _median_ms(frame, "ms_kernel_transport", reps)
        if not np.array_equal(frame.counts(), ref_counts):
            raise RuntimeError(f"tuning {cand} changed the latent counts")
        results.append({"knobs": cand, "ms_kernel_transport": ms})
        log("  transport " + " ".join(f"{k[len('transport_'):]} {v}" for k, v in zip(keys, values))
            + f": {ms:.2f} ms")
        if ms < best_ms:
            best_t, best_ms = cand, ms
    best_s, best_sms = None, float("inf")
    for rpl in SCAN_RAYS_PER_LANE:
        cand = dict(best_t, scan_rays_per_thread=rpl)
        engine.set_tuning(**cand)
        ms = _median_ms(frame, "ms_kernel_scan", reps)
        results.append({"knobs": cand, "ms_kernel_scan": ms})
        log(f"  scan rays per lane {rpl}: {ms:.2f} ms")
        if ms < best_sms:
            best_s, best_sms = cand, ms
    engine.set_tuning(**default)
    d_ms_t = next(r["ms_kernel_transport"] for r in results
                  if "ms_kernel_transport" in r and r["knobs"] == default)
    return {
        "workload": workload, "device": engine.adapter_name(), "crop": [w, h],
        "tuning": best_s, "ms_kernel_transport": best_ms, "ms_kernel_scan": best_sms,
        "default_ms_kernel_transport": d_ms_t,
        "measured": time.strftime("%Y-%m-%d"), "results": results,
    }


def main(argv=None):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import fgrain_amd as F

    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="frame24mp_trix_cubic")
    ap.add_argument("--tiles", default="8x6")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=None, help="output JSON (default: the tuning directory)")
    a = ap.parse_args(argv)
    tx, ty = (int(v) for v in a.tiles.split("x"))
    engine = F.GpuEngine.new()
    r = tune(engine, a.workload, (tx, ty), a.reps, log=lambda m: print(m, file=sys.stderr, flush=True))
    out = a.out or tuning_path(a.workload)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        json.dump(r, f, indent=1)
    print(json.dumps({k: r[k] for k in ("workload", "tuning", "ms_kernel_transport", "ms_kernel_scan",
                                        "default_ms_kernel_transport")}))


if __name__ == "__main__":
    main()
