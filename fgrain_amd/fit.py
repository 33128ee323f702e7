"""Inverse fit of an exposure map through the expected model.

The dye-layer pipeline of image_pipeline.rs:1035-1075 runs three independent B&W layers
(seeds +0, +10007, +20011) over one tiled image; each tile of a layer is its own emulsion
with the per-tile seed of image_pipeline.rs:430-434.  Here the target is the plain
(stochastic) per-tile forward's D_neg of a hidden exposure field; the fit runs Adam on
log2 exposure per pixel with gradients from fg_vjp (path replay) and is scored by the
plain forward again.  No reference counterpart: the reference has no gradients.
"""
from __future__ import annotations

import time

import numpy as np

from . import workloads as W
from .adjoint import ExpectedFilm
from .api import ScanMode, VirtualEmulsionLayer, develop_emulsion, expose_emulsion, scan_emulsion


def tile_seed(seed: int, idx: int, tx: int, ty: int) -> int:
    """Per-tile seed of the tiled renderer (image_pipeline.rs:430-434)."""
    return (seed + idx * 73856093 + tx * 19349663 + ty * 83492791) % (1 << 64)


class DyeFit:
    """Tiles x layers of expected films over one exposure map per layer."""

    def __init__(self, engine, cfg: dict, log=None):
        self.engine = engine
        self.cfg = cfg
        fr = cfg["frame"]
        self.width, self.height, self.tile = fr["width"], fr["height"], fr["tile"]
        self.weight = W.packet_weight(cfg)
        x = cfg["exposure"]
        self.lo, self.hi = x["lo"], x["hi"]
        # loss on the scanned density D_neg, or on the transmittance T = 10^(0.05 - D_neg)
        # (image_pipeline.rs:1006-1011 inverted), where a single stochastic target is an
        # unbiased sample of the expected model's value
        self.space = cfg.get("fit", {}).get("loss_space", "density")
        mode = ScanMode.NegativeTransmittance if self.space == "transmittance" else ScanMode.NegativeDensity
        offs = cfg.get("layer_seed_offsets", [0])[: cfg.get("layers", 1)]
        self.layers = []
        for off in offs:
            lseed = (cfg["seed"] + off) % (1 << 64)
            tiles = []
            idx = 0
            for ty, y0 in enumerate(range(0, self.height, self.tile)):
                for tx, x0 in enumerate(range(0, self.width, self.tile)):
                    tw, th = min(self.tile, self.width - x0), min(self.tile, self.height - y0)
                    s = tile_seed(lseed, idx, tx, ty)
                    em = VirtualEmulsionLayer.synthesize(W.emulsion_config(cfg, float(tw), float(th)), s)
                    film = ExpectedFilm(em, W.expected_model(cfg, tw, th, s, mode), engine)
                    tiles.append((x0, y0, tw, th, s, em, film))
                    idx += 1
            self.layers.append(tiles)
        if log:
            log(f"fit: {len(self.layers)} layers x {len(self.layers[0])} tiles, "
                f"{sum(len(t[5]) for L in self.layers for t in L)} crystals")

    # -- plain (stochastic) forward: fg_expose -> fg_develop -> fg_scan per tile --------
    def plain_forward(self, rel: list[np.ndarray], seed_shift: int = 0) -> list[np.ndarray]:
        """Each call exposes fresh film: exposures accumulate latent counts on an emulsion
        (crystal.rs:221-230), so the counts are cleared first."""
        out = []
        for layer, E in zip(self.layers, rel):
            D = np.zeros((self.height, self.width), np.float32)
            for (x0, y0, tw, th, s, em, film) in layer:
                em.set_state(counts=np.zeros(len(em), np.uint32))
                emap = np.ascontiguousarray(E[y0:y0 + th, x0:x0 + tw] * np.float32(self.weight))
                expose_emulsion(em, emap, W.exposure_config(self.cfg), s + 1 + seed_shift, self.engine)
                develop_emulsion(em, W.development_config(self.cfg), self.engine)
                D[y0:y0 + th, x0:x0 + tw], _ = scan_emulsion(em, W.scan_config(self.cfg, tw, th),
                                                             s + 2, self.engine)
            out.append(D)
        return out

    def hidden_field(self) -> list[np.ndarray]:
        f = self.cfg["fit"]
        return [W.smooth_field(self.width, self.height, f["hidden_seed"] + k, self.lo, self.hi,
                               self.cfg["exposure"].get("bumps", 8)) for k in range(len(self.layers))]

    # -- expected model: loss and gradient ----------------------------------------------
    def loss_grad(self, rel: list[np.ndarray], target: list[np.ndarray]):
        """L = mean over layers and pixels of (D_exp - target)^2; returns (L, dL/drel per
        layer, ms forward, ms vjp)."""
        n = sum(t.size for t in target)
        L = 0.0
        grads = []
        ms_f = ms_b = 0.0
        for layer, E, T in zip(self.layers, rel, target):
            g = np.zeros_like(E)
            for (x0, y0, tw, th, s, em, film) in layer:
                emap = np.ascontiguousarray(E[y0:y0 + th, x0:x0 + tw] * np.float32(self.weight))
                t0 = time.perf_counter()
                out, dose = film.forward(emap)
                t1 = time.perf_counter()
                r = out.astype(np.float64) - T[y0:y0 + th, x0:x0 + tw]
                L += float((r * r).sum())
                gE, _ = film.vjp(emap, (2.0 * r / n).astype(np.float32), dose=dose)
                t2 = time.perf_counter()
                ms_f += (t1 - t0) * 1e3
                ms_b += (t2 - t1) * 1e3
                g[y0:y0 + th, x0:x0 + tw] = gE * self.weight
            grads.append(g)
        return L / n, grads, ms_f, ms_b

    def run(self, steps: int | None = None, lr: float | None = None, log=None) -> dict:
        f = self.cfg["fit"]
        steps = f["steps"] if steps is None else steps
        lr = f["lr"] if lr is None else lr
        hidden = self.hidden_field()
        # warm-up: the first expected forward of a tile builds its expected-scan BVH and
        # touch graph on the device
        for layer in self.layers:
            for (x0, y0, tw, th, s, em, film) in layer:
                film.forward(np.full((th, tw), f["init_exposure"] * self.weight, np.float32))
        t = time.perf_counter()
        target_d = self.plain_forward(hidden)
        ms_target = (time.perf_counter() - t) * 1e3
        target = ([(10.0 ** (0.05 - d.astype(np.float64))).astype(np.float32) for d in target_d]
                  if self.space == "transmittance" else target_d)
        x = [np.full((self.height, self.width), np.log2(f["init_exposure"]), np.float64)
             for _ in self.layers]
        m = [np.zeros_like(a) for a in x]
        v = [np.zeros_like(a) for a in x]
        b1, b2, eps = 0.9, 0.999, 1e-8
        hist = []
        ms_f = ms_b = ms_step = 0.0
        for k in range(1, steps + 1):
            t0 = time.perf_counter()
            rel = [np.clip(2.0 ** a, 1e-4, 1e4).astype(np.float32) for a in x]
            L, gE, tf, tb = self.loss_grad(rel, target)
            for j in range(len(x)):
                gx = gE[j].astype(np.float64) * rel[j] * np.log(2.0)     # d/dlog2 E
                m[j] = b1 * m[j] + (1 - b1) * gx
                v[j] = b2 * v[j] + (1 - b2) * gx * gx
                x[j] -= lr * (m[j] / (1 - b1 ** k)) / (np.sqrt(v[j] / (1 - b2 ** k)) + eps)
            dt = (time.perf_counter() - t0) * 1e3
            ms_f += tf
            ms_b += tb
            ms_step += dt
            hist.append(L)
            if log and (k == 1 or k % 10 == 0 or k == steps):
                log(f"fit step {k}: loss {L:.4e}  {dt:.0f} ms (forward {tf:.0f}, adjoint {tb:.0f})")
        fitted = [np.clip(2.0 ** a, 1e-4, 1e4).astype(np.float32) for a in x]
        init = [np.full((self.height, self.width), f["init_exposure"], np.float32) for _ in self.layers]

        def rmse(a, b):
            return float(np.sqrt(np.mean([np.mean((p.astype(np.float64) - q) ** 2) for p, q in zip(a, b)])))
        plain_fit = self.plain_forward(fitted)
        plain_init = self.plain_forward(init)
        # a second photon realisation of the hidden field: the grain-noise floor
        plain_truth2 = self.plain_forward(hidden, seed_shift=7919)
        return {
            "steps": steps, "lr": lr, "loss_first": hist[0], "loss_last": hist[-1], "loss": hist,
            "ms_per_step": ms_step / steps, "ms_forward_per_step": ms_f / steps,
            "ms_adjoint_per_step": ms_b / steps, "adjoint_forward_ratio": ms_b / ms_f,
            "ms_target": ms_target,
            "loss_space": self.space,
            "rmse_plain_fit": rmse(plain_fit, target_d), "rmse_plain_init": rmse(plain_init, target_d),
            "rmse_plain_truth_resampled": rmse(plain_truth2, target_d),
            "rel_exposure_rmse_log2": float(np.sqrt(np.mean([np.mean((np.log2(a) - np.log2(b)) ** 2)
                                                            for a, b in zip(fitted, hidden)]))),
            "rel_exposure_rmse_log2_init": float(np.sqrt(np.mean([np.mean((np.log2(a) - np.log2(b)) ** 2)
                                                                 for a, b in zip(init, hidden)]))),
        }
