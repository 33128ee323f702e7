"""Synthetic workloads (committed JSON configs) and device-resident frames.

A workload names its sizes: ``tile320_trix_cubic`` (one 320x320 px tile) and
``frame24mp_trix_cubic`` (6000x4000 px as 247 tiles).  Exposures are synthetic,
generated from the seed in the config; no image files are read.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _lib as L
from ._lib import check, ptr
from .api import (DevelopmentConfig, EmulsionConfig, ExposureConfig, GelatinOptics, GpuEngine,
                  Morphology, ScanConfig, ScanMode)

CONFIG_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs")


def load_config(name: str) -> dict:
    with open(os.path.join(CONFIG_DIR, name + ".json")) as f:
        return json.load(f)


def emulsion_config(cfg: dict, width: float, height: float) -> EmulsionConfig:
    e = cfg["emulsion"]
    o = e["optics"]
    return EmulsionConfig(
        (width, height, e["depth"]), e["r_min"], e["mean_crystal_radius"], e["size_variance"],
        Morphology(e["morphology"]), e["tabular_aspect_ratio"],
        GelatinOptics(o["refractive_index"], o["absorption_coeff"], o["scattering_coeff"],
                      o["anisotropy_g"], o["antihalation_reflectance"], o["antihalation_density"]),
        e["base_sensitivity"], e["sensitivity_sigma"], e["min_radius"], e["min_half_thickness"],
        e["k_samples"])


def exposure_config(cfg: dict) -> ExposureConfig:
    x = cfg["exposure"]
    return ExposureConfig(0.0, x["illumination_angle_rad"], x["crystal_absorption_coeff"],
                          x["max_bounces"], x["photons_per_pixel"], x.get("base_thickness", 0.0))


def development_config(cfg: dict) -> DevelopmentConfig:
    d = cfg["development"]
    return DevelopmentConfig(d["threshold_photons"], d["filament_expansion"],
                             d["clumping_proximity_factor"], 0.85)


def scan_config(cfg: dict, w: int, h: int, mode=ScanMode.NegativeDensity) -> ScanConfig:
    s = cfg["scan"]
    return ScanConfig(w, h, s["samples_per_pixel"], mode, s["silver_opacity"], s["supersample"],
                      bool(s.get("otf", 0)))


def ramp_exposure(w: int, h: int, lo: float, hi: float) -> np.ndarray:
    """Linear horizontal ramp, first column lo, last column hi."""
    x = np.arange(w, dtype=np.float64)
    row = lo + (hi - lo) * x / max(w - 1, 1)
    return np.ascontiguousarray(np.broadcast_to(row.astype(np.float32), (h, w)))


def blob_exposure(w: int, h: int, seed: int, blobs=8, sigma=(200.0, 800.0), amp=(0.2, 0.8),
                  lo=0.01, hi=4.0) -> np.ndarray:
    """Seeded 8-bit luminance: vertical 0..1 ramp plus Gaussian blobs, clipped, quantised
    to 8 bits and mapped linearly to [lo, hi]."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(amp[0], amp[1], blobs)
    s = rng.uniform(sigma[0], sigma[1], blobs)
    cx = rng.uniform(0, w, blobs)
    cy = rng.uniform(0, h, blobs)
    y = np.arange(h, dtype=np.float64)[:, None]
    x = np.arange(w, dtype=np.float64)[None, :]
    lum = np.repeat(y / max(h - 1, 1), w, axis=1)
    for k in range(blobs):
        gx = np.exp(-((x - cx[k]) ** 2) / (2 * s[k] ** 2))
        gy = np.exp(-((y - cy[k]) ** 2) / (2 * s[k] ** 2))
        lum += a[k] * (gy * gx)
    lum = np.clip(lum, 0.0, 1.0)
    l8 = np.round(lum * 255.0)
    return (lo + (hi - lo) * l8 / 255.0).astype(np.float32)


def smooth_field(w: int, h: int, seed: int, lo: float, hi: float, bumps: int = 8) -> np.ndarray:
    """Seeded smooth field: a sum of Gaussian bumps (widths 10-35 % of the larger side)
    rescaled to [lo, hi]."""
    rng = np.random.default_rng(seed)
    y = np.arange(h, dtype=np.float64)[:, None]
    x = np.arange(w, dtype=np.float64)[None, :]
    f = np.zeros((h, w))
    for _ in range(bumps):
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        s = rng.uniform(0.10, 0.35) * max(w, h)
        f += rng.uniform(0.3, 1.0) * (np.exp(-((y - cy) ** 2) / (2 * s * s)) *
                                      np.exp(-((x - cx) ** 2) / (2 * s * s)))
    f = (f - f.min()) / max(f.max() - f.min(), 1e-300)
    return (lo + (hi - lo) * f).astype(np.float32)


def packet_weight(cfg: dict) -> float:
    """Photon weight per packet at relative exposure 1: photons per px^2 over packets per
    pixel (configs whose exposure names photons_per_px2; else 1)."""
    x = cfg["exposure"]
    d = x.get("photons_per_px2")
    return float(d) / x["photons_per_pixel"] if d else 1.0


def workload_exposure(cfg: dict) -> np.ndarray:
    fr, x = cfg["frame"], cfg["exposure"]
    if x["kind"] == "smooth_field":
        return smooth_field(fr["width"], fr["height"], cfg["seed"], x["lo"], x["hi"], x.get("bumps", 8))
    if x["kind"] == "ramp_x":
        return ramp_exposure(fr["width"], fr["height"], x["lo"], x["hi"])
    if x["kind"] == "blobs":
        return blob_exposure(fr["width"], fr["height"], cfg["seed"], x["blobs"], x["sigma"],
                             x["amplitude"], x["lo"], x["hi"])
    raise ValueError(x["kind"])


def density_multipliers(cfg: dict, n_tiles: int):
    """Seeded per-tile crystal-density multipliers U(low, high), or None."""
    dm = cfg.get("density_multiplier")
    if not dm:
        return None
    rng = np.random.default_rng(cfg["seed"] + 1)
    return rng.uniform(dm["low"], dm["high"], n_tiles).astype(np.float32)


def expected_model(cfg: dict, w: int, h: int, seed: int, mode=ScanMode.NegativeDensity):
    """ExpectedModel of one tile of the workload with tile seed `seed` (transport seed + 1,
    scan seed + 2, as the per-tile path)."""
    from .adjoint import ExpectedModel
    e = cfg["emulsion"]["optics"]
    optics = GelatinOptics(e["refractive_index"], e["absorption_coeff"], e["scattering_coeff"],
                           e["anisotropy_g"], e["antihalation_reflectance"], e["antihalation_density"])
    return ExpectedModel(exposure_config(cfg), optics, development_config(cfg), scan_config(cfg, w, h, mode),
                         expose_seed=(seed + 1) % (1 << 64), scan_seed=(seed + 2) % (1 << 64))


class Frame:
    """A device-resident frame of tiles (fg_frame_*): per-tile emulsions synthesised on
    the host once, then rendered any number of times on the GPU."""

    PLAIN = 0
    BATCHED = 1

    def __init__(self, engine: GpuEngine, cfg: dict, exposure: np.ndarray | None = None,
                 density_mult: np.ndarray | None = None, threads: int = 0,
                 frame: tuple | None = None, synthesis: str | None = None, overlap: int = 0):
        fr = cfg["frame"]
        w, h = (frame or (fr["width"], fr["height"]))
        self.width, self.height, self.tile = w, h, fr["tile"]
        self.seed = cfg["seed"]
        self.exposure = np.ascontiguousarray(
            exposure if exposure is not None else workload_exposure(cfg), np.float32)
        if self.exposure.shape != (h, w):
            raise ValueError("exposure map shape must be (height, width)")
        self.engine = engine
        d = L.FrameDescC()
        d.width, d.height, d.tile, d.threads = w, h, fr["tile"], threads
        d.pixel_size = 1.0
        d.emulsion = emulsion_config(cfg, 0.0, 0.0).c()
        d.exposure = exposure_config(cfg).c()
        d.development = development_config(cfg).c()
        d.scan = scan_config(cfg, 0, 0).c()
        d.seed = cfg["seed"]
        d.exposure_map = ptr(self.exposure, L._f32p)
        # "host": the reference's Bridson sampler; "device": GPU dart throwing
        self.synthesis = synthesis or fr.get("synthesis", "host")
        if self.synthesis not in ("host", "device"):
            raise ValueError("synthesis must be 'host' or 'device'")
        d.synthesis = 1 if self.synthesis == "device" else 0
        # overlapping tiles as the reference's tiled renderer lays them out
        self.overlap = int(overlap)
        d.overlap = self.overlap
        if density_mult is None:
            n_tiles = -(-w // fr["tile"]) * -(-h // fr["tile"])
            density_mult = density_multipliers(cfg, n_tiles)
        self._mult = None if density_mult is None else np.ascontiguousarray(density_mult, np.float32)
        d.density_mult = ptr(self._mult, L._f32p)
        self._h = C.c_void_p()
        check(L.lib().fg_frame_create(engine._h, C.byref(d), C.byref(self._h)))

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and L is not None and L._lib is not None:
            L._lib.fg_frame_destroy(h)

    @property
    def crystals(self) -> int:
        return int(L.lib().fg_frame_crystal_count(self._h))

    def render(self, mode: int = PLAIN) -> dict:
        st = L.FrameStatsC()
        check(L.lib().fg_frame_render(self.engine._h, self._h, mode, C.byref(st)))
        return {k: getattr(st, k) for k, _ in L.FrameStatsC._fields_}

    def read(self) -> np.ndarray:
        out = np.zeros((self.height, self.width), np.float32)
        check(L.lib().fg_frame_read(self.engine._h, self._h, ptr(out, L._f32p)))
        return out

    def development(self):
        """(developed u8, filament expansion f64, clump id i64) of the last render."""
        n = self.crystals
        d, f, c = np.zeros(n, np.uint8), np.zeros(n), np.zeros(n, np.int64)
        check(L.lib().fg_frame_read_development(self.engine._h, self._h, n, ptr(d, L._u8p),
                                                ptr(f, L._f64p), ptr(c, L._i64p)))
        return d, f, c

    def tile_data(self, k: int) -> dict:
        """Tile k's crystals (crystal order) and the device layout its transport reads:
        "nodes" (n_nodes x 8 u32 words, 32-byte threaded nodes), "records" (n x 12 words,
        48-byte records in leaf order)."""
        n, nn = C.c_int64(), C.c_int64()
        check(L.lib().fg_frame_get_tile(self.engine._h, self._h, k, C.byref(n), C.byref(nn), None, None, None, None,
                                        None, None, None))
        m = n.value
        out = dict(pos=np.zeros((m, 3)), dims=np.zeros((m, 3)), quat=np.zeros((m, 4)), morph=np.zeros(m, np.int32),
                   sens=np.zeros(m), nodes=np.zeros((nn.value, 8), np.uint32), records=np.zeros((m, 12), np.uint32))
        check(L.lib().fg_frame_get_tile(self.engine._h, self._h, k, None, None, ptr(out["pos"], L._f64p),
                                        ptr(out["dims"], L._f64p), ptr(out["quat"], L._f64p),
                                        ptr(out["morph"], L._i32p), ptr(out["sens"], L._f64p),
                                        out["nodes"].ctypes.data_as(C.c_void_p),
                                        out["records"].ctypes.data_as(C.c_void_p)))
        return out

    def tile_ranges(self):
        """(x0, y0, tw, th, seed) per tile, in frame order (image_pipeline.rs:397-434)."""
        out = []
        nx = -(-self.width // self.tile)
        ny = -(-self.height // self.tile)
        ov = getattr(self, "overlap", 0)
        for ty in range(ny):
            for tx in range(nx):
                idx = ty * nx + tx
                seed = (self.seed + idx * 73856093 + tx * 19349663 + ty * 83492791) % (1 << 64)
                x0, y0 = max(tx * self.tile - ov, 0), max(ty * self.tile - ov, 0)
                x1 = min(tx * self.tile + self.tile + ov, self.width)
                y1 = min(ty * self.tile + self.tile + ov, self.height)
                out.append((x0, y0, x1 - x0, y1 - y0, seed))
        return out

    def read_tiles(self) -> np.ndarray:
        """D_neg of every tile of the last render, concatenated in frame order."""
        n = C.c_int64()
        check(L.lib().fg_frame_read_tiles(self.engine._h, self._h, 0, None, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        check(L.lib().fg_frame_read_tiles(self.engine._h, self._h, n.value, ptr(out, L._f32p), None))
        return out

    def counts(self) -> np.ndarray:
        n = self.crystals
        out = np.zeros(n, np.uint32)
        check(L.lib().fg_frame_read_counts(self.engine._h, self._h, n, ptr(out, L._u32p)))
        return out
