"""Algorithmic bytes per photon path and per scanned pixel of a workload, counted by the
oracle running the reference algorithm (full infinite-ray traversal per bounce) on a
sample of the workload's tiles with the workload's seeds.  Writes roofline_bytes.json.

  bytes/path = 4 (exposure read) + 32 * nodes + 48 * crystal records + 8 * absorbing hits
  bytes/px   = rays/px * (32 * nodes + 48 * records per ray) + 4 (D write)

Per-path means do not depend on photons per pixel, so a reduced ppp is used."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))  # repository root

from fgrain_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main(name="frame24mp_trix_cubic", tiles=(0, 9, 18, 125, 133, 228, 246), ppp=8):
    cfg = W.load_config(name)
    e, o, x = cfg["emulsion"], cfg["emulsion"]["optics"], cfg["exposure"]
    exp = W.workload_exposure(cfg)
    fr = cfg["frame"]
    nx = -(-fr["width"] // fr["tile"])
    ny = -(-fr["height"] // fr["tile"])
    mult = W.density_multipliers(cfg, nx * ny)
    tot = dict(photons=0, nodes=0, prims=0, abs=0)
    for idx in tiles:
        tx, ty = idx % nx, idx // nx
        x0, y0 = tx * fr["tile"], ty * fr["tile"]
        tw, th = min(fr["tile"], fr["width"] - x0), min(fr["tile"], fr["height"] - y0)
        seed = (cfg["seed"] + idx * 73856093 + tx * 19349663 + ty * 83492791) % (1 << 64)
        r_min = e["r_min"] / (float(mult[idx]) ** (1.0 / 3.0) if mult is not None else 1.0)
        cs = O.synthesize(W=tw, H=th, D=e["depth"], r_min=r_min, mean_radius=e["mean_crystal_radius"],
                          size_sigma=e["size_variance"], morph=e["morphology"], aspect=e["tabular_aspect_ratio"],
                          base_sens=e["base_sensitivity"], sens_sigma=e["sensitivity_sigma"],
                          min_radius=e["min_radius"], min_half_thickness=e["min_half_thickness"], seed=seed)
        _, st = O.expose(cs, np.ascontiguousarray(exp[y0:y0 + th, x0:x0 + tw]), W=tw, H=th, D=e["depth"],
                         mu_a=o["absorption_coeff"], mu_s=o["scattering_coeff"], g=o["anisotropy_g"],
                         n_refr=o["refractive_index"], r_ahu=o["antihalation_reflectance"],
                         od_ahu=o["antihalation_density"], mu_abs=x["crystal_absorption_coeff"],
                         theta_max=x["illumination_angle_rad"], max_bounces=x["max_bounces"],
                         seed=seed + 1, sampler=1, ppp=ppp, base_thickness=x.get("base_thickness", 0.0))
        tot["photons"] += st["launched"]
        tot["nodes"] += st["nodes_tested"]
        tot["prims"] += st["prims_tested"]
        tot["abs"] += st["abs_events"]
        print(idx, tw, th, cs.n, st["nodes_tested"] / st["launched"], flush=True)
    p = tot["photons"]
    per = dict(nodes=tot["nodes"] / p, prims=tot["prims"] / p, abs_events=tot["abs"] / p)
    per["bytes"] = 4 + 32 * per["nodes"] + 48 * per["prims"] + 8 * per["abs_events"]
    out_path = os.path.join(HERE, "roofline_bytes.json")
    data = json.load(open(out_path)) if os.path.exists(out_path) else {}
    data[name] = {"transport_per_path": per, "sample_tiles": list(tiles), "photons_per_pixel": ppp,
                  "photons": p, "definition": __doc__.split("\n\n")[0]}
    json.dump(data, open(out_path, "w"), indent=1)
    print(json.dumps(data[name], indent=1))


if __name__ == "__main__":
    main(*sys.argv[1:2])
