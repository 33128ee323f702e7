"""Full-frame black-and-white film render with every stage on the MI355X
(render_film_black_and_white_tiled, GuyNguyen/fgrain src/image_pipeline.rs:299-904).

    report, out = render_film_black_and_white_tiled(image, FilmRenderOptions(...))

``image`` is an (h, w) gray or (h, w, 3) RGB uint8 array at output resolution; ``out`` is
the (h, w) uint8 print.  The stages -- panchromatic luma (:259-296), camera and gelatin
optics (:133-252), the overlapping tiles as one batched frame (:380-592), Hann blend,
low-frequency separation, sensitometric compositing, scanner OTF and 8-bit output
(:593-891) -- run in libfgrain_hip.so (fg_render_film_bw); the stage entry points are
exposed for tests and for callers that bring their own tiles.

Resampling to another output resolution and supersampling (:306-343) are the ``image``
crate's filters; the colour path (:1017-1085) renders three of these frames.  Both stay
with the caller here and are rejected with NotImplementedError.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import check, ptr
from .api import GpuEngine, get_gpu_engine


@dataclass
class FilmRenderOptions:
    """FilmRenderOptions (image_pipeline.rs:20-92) and its defaults, plus the device
    synthesis choice and frame mode of this engine."""
    preset: str = "tri-x"
    grain_size_px: float = 2.2
    grain_strength: float = 1.0
    exposure: float = 1.0
    color: bool = False
    output_resolution: tuple | None = None
    tile_size: int = 320
    seed: int = 42
    gpu: bool = True          # the render always runs on the device
    auto_scale: bool = False
    scale: float = 1.0
    reference_res: int = 6000
    supersample: int = 1
    push: float = 0.0
    eberhard: float = 1.0     # unused by the reference's black-and-white path (halo = 0)
    scanner_otf: bool = True
    synthesis: str = "host"   # "host": the reference's Bridson sampler; "device": dart throwing
    frame_mode: str = "batched"
    threads: int = 0

    def c(self) -> L.FilmOptionsC:
        o = L.FilmOptionsC()
        check(L.lib().fg_film_options_default(C.byref(o)))
        name = self.preset.encode()
        if len(name) >= 32:
            raise ValueError("preset name longer than 31 bytes")
        o.preset = name
        o.grain_size_px = self.grain_size_px
        o.grain_strength = self.grain_strength
        o.exposure = self.exposure
        o.scale = self.scale
        o.push = self.push
        o.seed = self.seed % (1 << 64)
        o.tile_size = self.tile_size
        o.reference_res = self.reference_res
        o.auto_scale = int(bool(self.auto_scale))
        o.scanner_otf = int(bool(self.scanner_otf))
        if self.synthesis not in ("host", "device"):
            raise ValueError("synthesis must be 'host' or 'device'")
        o.synthesis = 1 if self.synthesis == "device" else 0
        if self.frame_mode not in ("batched", "plain"):
            raise ValueError("frame_mode must be 'batched' or 'plain'")
        o.frame_mode = 1 if self.frame_mode == "batched" else 0
        o.threads = self.threads
        return o


@dataclass
class ImageProcessReport:
    """ImageProcessReport (image_pipeline.rs:95-105) with real development counts, the
    compositing statistics and per-stage times (ms)."""
    input_resolution: tuple
    output_resolution: tuple
    total_tiles: int
    total_crystals: int
    developed_crystals: int
    total_clumps: int
    elapsed_seconds: float
    photons: int = 0
    rays: int = 0
    stats: dict = field(default_factory=dict)
    times: dict = field(default_factory=dict)


def _engine(engine):
    return engine if engine is not None else get_gpu_engine()


def _check_supported(options: FilmRenderOptions, w: int, h: int):
    if options.color:
        raise NotImplementedError("the colour path renders three black-and-white frames; "
                                  "call render_film_black_and_white_tiled per dye layer")
    if options.output_resolution not in (None, (w, h)):
        raise NotImplementedError("resample the image to the output resolution first")
    if int(options.supersample) > 1:
        raise NotImplementedError("supersampling re-renders at a larger size and resamples; "
                                  "render at the larger size and resample instead")


def film_plan(options: FilmRenderOptions, width: int, height: int) -> dict:
    """Derived sizes (image_pipeline.rs:346-385, :651-658); host only."""
    p = L.FilmPlanC()
    check(L.lib().fg_film_plan_for(C.byref(options.c()), width, height, C.byref(p)))
    return {k: getattr(p, k) for k, _ in L.FilmPlanC._fields_ if k != "pad_"}


def render_film_black_and_white_tiled(image, options: FilmRenderOptions | None = None,
                                      engine: GpuEngine | None = None):
    """(report, uint8 image) of the whole render (fg_render_film_bw)."""
    options = options or FilmRenderOptions()
    img = np.ascontiguousarray(image, np.uint8)
    if img.ndim == 3 and img.shape[2] == 3:
        ch = 3
    elif img.ndim == 2:
        ch = 1
    else:
        raise ValueError("image must be (h, w) gray or (h, w, 3) RGB uint8")
    h, w = img.shape[:2]
    _check_supported(options, w, h)
    eng = _engine(engine)
    out = np.zeros((h, w), np.uint8)
    r = L.FilmReportC()
    check(L.lib().fg_render_film_bw(eng._h, ptr(img, L._u8p), w, h, ch, C.byref(options.c()),
                                    ptr(out, L._u8p), C.byref(r)))
    rep = ImageProcessReport(
        (r.in_w, r.in_h), (r.out_w, r.out_h), r.total_tiles, r.total_crystals, r.developed_crystals,
        r.total_clumps, r.elapsed_seconds, r.photons, r.rays,
        {k: getattr(r, k) for k in ("high_mean", "high_std", "norm_factor")},
        {k: getattr(r, k) for k in ("ms_luma_optics", "ms_synthesis", "ms_render", "ms_composite",
                                    "ms_kernel_optics", "ms_kernel_composite", "ms_kernel_transport",
                                    "ms_kernel_scan")})
    return rep, out


def film_luma(rgb, preset="tri-x", engine: GpuEngine | None = None, timing: dict | None = None) -> np.ndarray:
    """convert_to_panchromatic_luma of an (h, w, 3) uint8 image on the device;
    timing["ms_kernel"] receives the kernel time when a dict is given."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    if rgb.ndim != 3 or rgb.shape[2] != 3:
        raise ValueError("rgb must be (h, w, 3) uint8")
    h, w = rgb.shape[:2]
    out = np.zeros((h, w), np.uint8)
    ms = C.c_float()
    check(L.lib().fg_film_luma(_engine(engine)._h, ptr(rgb, L._u8p), w, h, preset.encode(),
                               ptr(out, L._u8p), C.byref(ms)))
    if timing is not None:
        timing["ms_kernel"] = ms.value
    return out


def film_optics(gray, engine: GpuEngine | None = None, timing: dict | None = None) -> np.ndarray:
    """apply_camera_and_gelatin_optics: h_effective (f32) of an (h, w) uint8 image;
    timing["ms_kernel"] receives the gate + optics kernel time when a dict is given."""
    g = np.ascontiguousarray(gray, np.uint8)
    h, w = g.shape
    out = np.zeros((h, w), np.float32)
    ms = C.c_float()
    check(L.lib().fg_film_optics(_engine(engine)._h, ptr(g, L._u8p), w, h, ptr(out, L._f32p), C.byref(ms)))
    if timing is not None:
        timing["ms_kernel"] = ms.value
    return out


def film_composite(gray, h_eff, tiles, options: FilmRenderOptions | None = None,
                   engine: GpuEngine | None = None, intermediates: bool = False):
    """Compositing of rendered tiles (image_pipeline.rs:593-891) for the options' tile
    layout: tiles = every tile's D_neg concatenated in frame order.  Returns (uint8 image,
    stats[, {"d_norm", "d_high", "grain"}])."""
    options = options or FilmRenderOptions()
    g = np.ascontiguousarray(gray, np.uint8)
    h, w = g.shape
    he = np.ascontiguousarray(h_eff, np.float32)
    t = np.ascontiguousarray(tiles, np.float32).reshape(-1)
    if he.shape != (h, w):
        raise ValueError("h_eff must match the image")
    out = np.zeros((h, w), np.uint8)
    bufs = {k: np.zeros((h, w), np.float32) for k in ("d_norm", "d_high", "grain")} if intermediates else {}
    st = L.CompositeStatsC()
    check(L.lib().fg_film_composite(_engine(engine)._h, C.byref(options.c()), w, h, ptr(g, L._u8p),
                                    ptr(he, L._f32p), ptr(t, L._f32p), t.size, ptr(out, L._u8p),
                                    *[ptr(bufs.get(k), L._f32p) for k in ("d_norm", "d_high", "grain")],
                                    C.byref(st)))
    stats = {k: getattr(st, k) for k, _ in L.CompositeStatsC._fields_ if not k.startswith("pad")}
    return (out, stats, bufs) if intermediates else (out, stats)


__all__ = ["FilmRenderOptions", "ImageProcessReport", "film_plan", "render_film_black_and_white_tiled",
           "film_luma", "film_optics", "film_composite"]
