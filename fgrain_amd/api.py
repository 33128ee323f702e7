"""Host-side mirror of fgrain's library API over the HIP engine.

Names, argument meaning and error behaviour follow the reference crate
(GuyNguyen/fgrain src/lib.rs:10-34): ``EmulsionConfig`` presets,
``VirtualEmulsionLayer.synthesize``, ``expose_emulsion``, ``develop_emulsion``,
``scan_emulsion``, ``GpuEngine`` / ``get_gpu_engine`` / ``render_tile_all_gpu`` and
``FilmSimulation.run``.  Every stage runs through libfgrain_hip.so; there is no CPU
path in this package.

Differences from the Rust signatures, all forced by the device boundary:
* ``expose_emulsion`` takes an exposure *map* (looked up with the reference tile
  path's nearest-pixel rule, image_pipeline.rs:949-961) instead of a closure;
* ``develop_emulsion`` needs no BVH argument (the emulsion owns it);
* GPU failures raise ``GpuError`` (status + message) instead of returning ``Err``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np

from . import _lib as L
from ._lib import GpuError, check, ptr


class Morphology(IntEnum):
    """crystal.rs:6-18"""
    Cubic = 0
    Tabular = 1
    Octahedral = 2


@dataclass
class GelatinOptics:
    """optics.rs:41-73 (defaults of GelatinOptics::default)."""
    refractive_index: float = 1.534
    absorption_coeff: float = 0.002
    scattering_coeff: float = 0.045
    anisotropy_g: float = 0.82
    antihalation_reflectance: float = 0.12
    antihalation_density: float = 0.85

    def extinction_coeff(self) -> float:
        return self.absorption_coeff + self.scattering_coeff

    def albedo(self) -> float:
        mt = self.extinction_coeff()
        return self.scattering_coeff / mt if mt > 0 else 0.0

    def c(self) -> L.Optics:
        return L.Optics(self.refractive_index, self.absorption_coeff, self.scattering_coeff,
                        self.anisotropy_g, self.antihalation_reflectance,
                        self.antihalation_density)


@dataclass
class EmulsionConfig:
    """EmulsionConfig (emulsion.rs:14-38) with the presets of emulsion.rs:46-136."""
    dimensions: tuple = (20.0, 20.0, 14.0)
    r_min: float = 1.2
    mean_crystal_radius: float = 0.55
    size_variance: float = 0.40
    morphology: Morphology = Morphology.Cubic
    tabular_aspect_ratio: float = 1.0
    optics: GelatinOptics = field(default_factory=GelatinOptics)
    base_sensitivity: float = 1.2
    sensitivity_sigma: float = 0.15
    min_radius: float = 0.15
    min_half_thickness: float = 0.05
    k_samples: int = 30

    @classmethod
    def tri_x_400(cls, width, height):
        return cls((width, height, 14.0), 1.2, 0.55, 0.40, Morphology.Cubic, 1.0,
                   GelatinOptics(1.534, 0.003, 0.050, 0.80, 0.14, 0.80), 1.2)

    @classmethod
    def tri_x_1600(cls, width, height):
        return cls((width, height, 14.5), 1.1, 0.52, 0.38, Morphology.Cubic, 1.0,
                   GelatinOptics(1.534, 0.0032, 0.052, 0.80, 0.14, 0.80), 1.35)

    @classmethod
    def t_max_100(cls, width, height):
        return cls((width, height, 8.0), 0.9, 0.40, 0.20, Morphology.Tabular, 6.5,
                   GelatinOptics(1.534, 0.002, 0.035, 0.85, 0.08, 1.10), 0.8)

    @classmethod
    def hp5_plus(cls, width, height):
        return cls((width, height, 12.5), 1.1, 0.50, 0.30, Morphology.Octahedral, 1.0,
                   GelatinOptics(1.534, 0.0025, 0.045, 0.82, 0.12, 0.85), 1.1)

    def c(self) -> L.EmulsionConfigC:
        w, h, d = self.dimensions
        return L.EmulsionConfigC(w, h, d, self.r_min, self.mean_crystal_radius,
                                 self.size_variance, self.tabular_aspect_ratio,
                                 self.base_sensitivity, self.sensitivity_sigma,
                                 self.min_radius, self.min_half_thickness,
                                 int(self.morphology), self.k_samples, self.optics.c())


class VirtualEmulsionLayer:
    """VirtualEmulsionLayer (emulsion.rs:141-226): crystals + BVH, host-owned."""

    def __init__(self, handle, config: EmulsionConfig):
        self._h = handle
        self.config = config

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and L is not None and L._lib is not None:
            L._lib.fg_emulsion_destroy(h)

    @classmethod
    def synthesize(cls, config: EmulsionConfig, seed: int):
        h = C.c_void_p()
        cfg = config.c()
        check(L.lib().fg_emulsion_synthesize(C.byref(cfg), seed & (2**64 - 1), C.byref(h)))
        return cls(h, config)

    @classmethod
    def synthesize_device(cls, config: EmulsionConfig, seed: int, engine: "GpuEngine | None" = None):
        """Device synthesis (no reference counterpart): Poisson-disk centres by parallel
        dart throwing on the GPU, the reference's per-crystal draws and BVH on the host."""
        e = _engine(engine)
        h = C.c_void_p()
        cfg = config.c()
        check(L.lib().fg_emulsion_synthesize_device(e._h, C.byref(cfg), seed & (2**64 - 1), C.byref(h)))
        return cls(h, config)

    @classmethod
    def from_crystals(cls, config: EmulsionConfig, pos, dims, quat, morph, sens):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        n = pos.shape[0]
        dims = np.ascontiguousarray(dims, np.float64).reshape(n, 3)
        quat = np.ascontiguousarray(quat, np.float64).reshape(n, 4)
        morph = np.ascontiguousarray(np.broadcast_to(morph, (n,)), np.int32)
        sens = np.ascontiguousarray(np.broadcast_to(sens, (n,)), np.float64)
        h = C.c_void_p()
        cfg = config.c()
        check(L.lib().fg_emulsion_from_crystals(
            C.byref(cfg), n, ptr(pos, L._f64p), ptr(dims, L._f64p), ptr(quat, L._f64p),
            ptr(morph, L._i32p), ptr(sens, L._f64p), C.byref(h)))
        return cls(h, config)

    def __len__(self):
        return int(L.lib().fg_emulsion_count(self._h))

    def crystals(self) -> dict:
        n = len(self)
        out = dict(pos=np.zeros((n, 3)), dims=np.zeros((n, 3)), quat=np.zeros((n, 4)),
                   morph=np.zeros(n, np.int32), sens=np.zeros(n))
        check(L.lib().fg_emulsion_get_crystals(
            self._h, ptr(out["pos"], L._f64p), ptr(out["dims"], L._f64p),
            ptr(out["quat"], L._f64p), ptr(out["morph"], L._i32p), ptr(out["sens"], L._f64p)))
        return out

    def bvh(self) -> dict:
        """BVH::build layout (bvh.rs:129-151)."""
        nn = C.c_int64()
        check(L.lib().fg_emulsion_get_bvh(self._h, C.byref(nn), None, None, None, None, None,
                                          None, None))
        k, n = nn.value, len(self)
        o = dict(bmin=np.zeros((k, 3)), bmax=np.zeros((k, 3)), left=np.zeros(k, np.uint32),
                 right=np.zeros(k, np.uint32), offset=np.zeros(k, np.uint32),
                 count=np.zeros(k, np.uint32), indices=np.zeros(n, np.uint32))
        check(L.lib().fg_emulsion_get_bvh(
            self._h, C.byref(nn), ptr(o["bmin"], L._f64p), ptr(o["bmax"], L._f64p),
            ptr(o["left"], L._u32p), ptr(o["right"], L._u32p), ptr(o["offset"], L._u32p),
            ptr(o["count"], L._u32p), ptr(o["indices"], L._u32p)))
        return o

    def device_layout(self):
        """(nodes, records) as the GPU reads them: 32-byte threaded nodes, 48-byte records."""
        nn = C.c_int64()
        check(L.lib().fg_emulsion_get_bvh(self._h, C.byref(nn), None, None, None, None, None,
                                          None, None))
        nodes = np.zeros(nn.value * 8, np.uint32)
        recs = np.zeros(len(self) * 12, np.uint32)
        check(L.lib().fg_emulsion_get_device_layout(self._h, nodes.ctypes.data_as(C.c_void_p),
                                                    recs.ctypes.data_as(C.c_void_p)))
        return nodes.reshape(-1, 8), recs.reshape(-1, 12)

    def state(self) -> dict:
        n = len(self)
        o = dict(counts=np.zeros(n, np.uint32), developed=np.zeros(n, np.uint8),
                 fe=np.zeros(n), clump_id=np.zeros(n, np.int64))
        check(L.lib().fg_emulsion_get_state(self._h, ptr(o["counts"], L._u32p),
                                            ptr(o["developed"], L._u8p), ptr(o["fe"], L._f64p),
                                            ptr(o["clump_id"], L._i64p)))
        return o

    def absorbed_photons(self) -> np.ndarray:
        """HalideCrystal::absorbed_photons for every crystal (crystal.rs:226-230)."""
        return self.state()["counts"]

    def set_state(self, counts=None, developed=None, fe=None):
        n = len(self)
        c = None if counts is None else np.ascontiguousarray(counts, np.uint32).reshape(n)
        d = None if developed is None else np.ascontiguousarray(developed, np.uint8).reshape(n)
        f = None if fe is None else np.ascontiguousarray(fe, np.float64).reshape(n)
        check(L.lib().fg_emulsion_set_state(self._h, ptr(c, L._u32p), ptr(d, L._u8p),
                                            ptr(f, L._f64p)))


@dataclass
class ExposureConfig:
    """ExposureConfig (transport.rs:15-39); photons_per_pixel > 0 selects stratified
    sampling over the exposure map's pixels."""
    photons_per_sq_um: float = 50.0
    illumination_angle_rad: float = 0.15
    crystal_absorption_coeff: float = 1.2
    max_bounces: int = 64
    photons_per_pixel: int = 0
    base_thickness: float = 0.0

    def c(self):
        return L.ExposureConfigC(self.photons_per_sq_um, self.illumination_angle_rad,
                                 self.crystal_absorption_coeff, self.max_bounces,
                                 self.photons_per_pixel, self.base_thickness)


@dataclass
class DevelopmentConfig:
    """DevelopmentConfig (development.rs:12-38)."""
    threshold_photons: int = 3
    filament_expansion: float = 1.8
    clumping_proximity_factor: float = 1.25
    developer_contrast: float = 0.75

    def c(self):
        return L.DevelopmentConfigC(self.threshold_photons, 0, self.filament_expansion,
                                    self.clumping_proximity_factor, self.developer_contrast)


@dataclass
class DevelopmentReport:
    """DevelopmentReport (development.rs:42-55)."""
    total_crystals: int = 0
    developed_crystals: int = 0
    fixed_crystals: int = 0
    total_clumps: int = 0
    max_clump_size: int = 0
    avg_clump_size: float = 0.0


class ScanMode(IntEnum):
    """ScanMode (transmittance.rs:17-28) plus the tile path's D_neg."""
    NegativeTransmittance = 0
    PositivePrint = 1
    OpticalDensity = 2
    NegativeDensity = 3


@dataclass
class ScanConfig:
    """ScanConfig (transmittance.rs:32-60) + aperture supersampling + scanner OTF."""
    width_px: int = 512
    height_px: int = 512
    samples_per_pixel: int = 16
    mode: ScanMode = ScanMode.PositivePrint
    silver_opacity: float = 25.0
    supersample: int = 1
    otf: bool = False

    def c(self):
        return L.ScanConfigC(self.width_px, self.height_px, self.samples_per_pixel,
                             int(self.mode), self.silver_opacity, self.supersample,
                             1 if self.otf else 0)


class GpuEngine:
    """GpuEngine (gpu.rs:155-170) over the HIP device."""

    def __init__(self, handle, owned=True):
        self._h = handle
        self._owned = owned

    @classmethod
    def new(cls):
        """GpuEngine::new; raises GpuError(AdapterNotFound) without a HIP device."""
        h = C.c_void_p()
        check(L.lib().fg_engine_create(C.byref(h)))
        return cls(h)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and getattr(self, "_owned", False) and L is not None and L._lib is not None:
            L._lib.fg_engine_destroy(h)

    def adapter_name(self) -> str:
        return L.lib().fg_engine_adapter_name(self._h).decode()

    def set_tuning(self, transport_kernel=0, transport_chunk=64, transport_wait_pct=80,
                   scan_rays_per_thread=0, transport_waves=6, transport_prereduce=0):
        """Launch parameters (speed only; results do not depend on them)."""
        t = L.TuningC(transport_kernel, transport_chunk, transport_wait_pct, scan_rays_per_thread,
                      transport_waves, transport_prereduce)
        check(L.lib().fg_engine_set_tuning(self._h, C.byref(t)))

    def tuning(self) -> dict:
        t = L.TuningC()
        check(L.lib().fg_engine_get_tuning(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in L.TuningC._fields_}

    def render_tile_all_gpu(self, tile_img, params: "GpuTileParams") -> np.ndarray:
        """GpuEngine::render_tile_all_gpu (gpu.rs:491-938): D_neg of one tile."""
        p = params.c
        n = p.width_px * p.height_px
        if n == 0:
            return np.zeros(0, np.float32)
        img = np.ascontiguousarray(tile_img, np.float32).reshape(-1)
        if img.size != n:
            raise ValueError("tile_img must hold width_px * height_px values")
        out = np.zeros(n, np.float32)
        st = L.TileStatsC()
        check(L.lib().fg_render_tile_all_gpu(self._h, ptr(img, L._f32p), C.byref(p),
                                             ptr(out, L._f32p), C.byref(st)))
        self.last_tile_stats = {k: getattr(st, k) for k, _ in L.TileStatsC._fields_}
        return out

    def render_tile_delta_d(self, developed_crystals: dict, params: "GpuTileParams",
                            with_dneg: bool = False):
        """GpuEngine::render_tile_delta_d (gpu.rs:941-1228): the developed crystals
        (dict of arrays "pos" (n,3), "dims" (n,3) semi-axes, "quat" (n,4) (w,i,j,k),
        "morph" (n,), "fe" (n,) filament expansion) scanned as scan_emulsion does, then
        D_neg minus its 17-tap blur (gaussian.wgsl).  Returns width_px * height_px values (and D_neg
        itself with with_dneg)."""
        p = params.c
        n_px = p.width_px * p.height_px
        if n_px == 0:
            z = np.zeros(0, np.float32)
            return (z, z.copy()) if with_dneg else z
        pos = np.ascontiguousarray(developed_crystals["pos"], np.float64).reshape(-1, 3)
        n = pos.shape[0]
        dims = np.ascontiguousarray(developed_crystals["dims"], np.float64).reshape(n, 3)
        quat = np.ascontiguousarray(developed_crystals["quat"], np.float64).reshape(n, 4)
        morph = np.ascontiguousarray(developed_crystals["morph"], np.int32).reshape(n)
        fe = np.ascontiguousarray(developed_crystals["fe"], np.float64).reshape(n)
        out = np.zeros(n_px, np.float32)
        dneg = np.zeros(n_px, np.float32) if with_dneg else None
        check(L.lib().fg_render_tile_delta_d(self._h, n, ptr(pos, L._f64p), ptr(dims, L._f64p),
                                             ptr(quat, L._f64p), ptr(morph, L._i32p), ptr(fe, L._f64p),
                                             C.byref(p), ptr(out, L._f32p), ptr(dneg, L._f32p)))
        return (out, dneg) if with_dneg else out

    def traverse_all_hits(self, emulsion: VirtualEmulsionLayer, origins, dirs, max_hits=256):
        """BVH::intersect_ray_all on the GPU: (n_hits, idx, t0, t1) in DFS callback order."""
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = o.shape[0]
        nh = np.zeros(n, np.int32)
        idx = np.zeros((n, max_hits), np.int32)
        t0 = np.zeros((n, max_hits), np.float32)
        t1 = np.zeros((n, max_hits), np.float32)
        check(L.lib().fg_traverse_all_hits(self._h, emulsion._h, n, ptr(o, L._f32p), ptr(d, L._f32p),
                                           max_hits, ptr(nh, L._i32p), ptr(idx, L._i32p),
                                           ptr(t0, L._f32p), ptr(t1, L._f32p)))
        return nh, idx, t0, t1

    def first_hit(self, emulsion: VirtualEmulsionLayer, origins, dirs):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = o.shape[0]
        idx = np.zeros(n, np.int32)
        t0 = np.zeros(n, np.float32)
        t1 = np.zeros(n, np.float32)
        check(L.lib().fg_traverse_first_hit(self._h, emulsion._h, n, ptr(o, L._f32p),
                                            ptr(d, L._f32p), ptr(idx, L._i32p),
                                            ptr(t0, L._f32p), ptr(t1, L._f32p)))
        return idx, t0, t1


_GLOBAL = None


def get_gpu_engine():
    """get_gpu_engine (gpu.rs:9-14): the process-wide engine, or None when its creation
    failed (the failure is remembered)."""
    global _GLOBAL
    if _GLOBAL is None:
        h = C.c_void_p()
        if L.lib().fg_get_engine(C.byref(h)) != L.FG_OK:
            return None
        _GLOBAL = GpuEngine(h, owned=False)
    return _GLOBAL


def _engine(engine):
    if engine is not None:
        return engine
    e = get_gpu_engine()
    if e is None:
        check(L.lib().fg_get_engine(C.byref(C.c_void_p())))
    return e


def expose_emulsion(emulsion: VirtualEmulsionLayer, exposure_map, config: ExposureConfig,
                    seed: int, engine: GpuEngine | None = None) -> dict:
    """expose_emulsion (transport.rs:48-188) on the GPU; returns transport statistics."""
    e = _engine(engine)
    if exposure_map is None:
        em, tw, th = None, 0, 0
    else:
        em = np.ascontiguousarray(exposure_map, np.float32)
        th, tw = em.shape
    cfg = config.c()
    st = L.ExposeStatsC()
    check(L.lib().fg_expose(e._h, emulsion._h, ptr(em, L._f32p), tw, th, C.byref(cfg),
                            seed & (2**64 - 1), C.byref(st)))
    return {k: getattr(st, k) for k, _ in L.ExposeStatsC._fields_ if k != "pad_"}


def develop_emulsion(emulsion: VirtualEmulsionLayer, config: DevelopmentConfig,
                     engine: GpuEngine | None = None) -> DevelopmentReport:
    """develop_emulsion (development.rs:59-188)."""
    e = _engine(engine)
    cfg = config.c()
    rep = L.DevelopmentReportC()
    check(L.lib().fg_develop(e._h, emulsion._h, C.byref(cfg), C.byref(rep)))
    return DevelopmentReport(rep.total_crystals, rep.developed_crystals, rep.fixed_crystals,
                             rep.total_clumps, rep.max_clump_size, rep.avg_clump_size)


def scan_emulsion(emulsion: VirtualEmulsionLayer, config: ScanConfig, seed: int,
                  engine: GpuEngine | None = None):
    """scan_emulsion (transmittance.rs:65-171) -> (buffer (h, w) f32, 8-bit image)."""
    e = _engine(engine)
    out = np.zeros((config.height_px, config.width_px), np.float32)
    cfg = config.c()
    check(L.lib().fg_scan(e._h, emulsion._h, C.byref(cfg), seed & (2**64 - 1), ptr(out, L._f32p)))
    img = np.round(np.clip(out, 0.0, 1.0) * 255.0).astype(np.uint8)
    return out, img


class GpuTileParams:
    """GpuTileParams (gpu.rs:18-37) plus the preset values of the CPU tile path."""

    def __init__(self, **kw):
        self.c = L.TileParamsC()
        for k, v in kw.items():
            setattr(self.c, k, v)

    @classmethod
    def for_preset(cls, preset: str, grain_size_px: float, tw: int, th: int,
                   exposure: float = 1.0, push: float = 0.0, tile_seed: int = 42):
        p = cls()
        check(L.lib().fg_tile_params_init(C.byref(p.c), preset.encode(), grain_size_px, tw, th,
                                          exposure, push, tile_seed & (2**64 - 1)))
        return p


@dataclass
class FilmSimulation:
    """FilmSimulation (lib.rs:38-111): synthesize(seed) -> expose(seed+1) ->
    develop -> scan(seed+2)."""
    emulsion_config: EmulsionConfig
    exposure_config: ExposureConfig
    development_config: DevelopmentConfig
    scan_config: ScanConfig

    @classmethod
    def tri_x_400(cls, width_um, height_um, width_px, height_px):
        return cls(EmulsionConfig.tri_x_400(width_um, height_um), ExposureConfig(),
                   DevelopmentConfig(), ScanConfig(width_px, height_px, 16,
                                                   ScanMode.PositivePrint, 25.0))

    @classmethod
    def t_max_100(cls, width_um, height_um, width_px, height_px):
        return cls(EmulsionConfig.t_max_100(width_um, height_um), ExposureConfig(),
                   DevelopmentConfig(), ScanConfig(width_px, height_px, 16,
                                                   ScanMode.PositivePrint, 25.0))

    def run(self, exposure_map, seed: int, engine: GpuEngine | None = None):
        em = VirtualEmulsionLayer.synthesize(self.emulsion_config, seed)
        expose_emulsion(em, exposure_map, self.exposure_config, seed + 1, engine)
        report = develop_emulsion(em, self.development_config, engine)
        _, image = scan_emulsion(em, self.scan_config, seed + 2, engine)
        return report, image


__all__ = [
    "Morphology", "GelatinOptics", "EmulsionConfig", "VirtualEmulsionLayer", "ExposureConfig",
    "DevelopmentConfig", "DevelopmentReport", "ScanMode", "ScanConfig", "GpuEngine",
    "GpuError", "get_gpu_engine", "expose_emulsion", "develop_emulsion", "scan_emulsion",
    "GpuTileParams", "FilmSimulation",
]
