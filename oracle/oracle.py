"""ORACLE -- test infrastructure only.

ctypes front-end of ``liboracle.so``, the plain-C CPU restatement of the reference's
photon-transport + development + scan path (GuyNguyen/fgrain).  Only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import this module,
and only to check or to time the CPU baseline; the product never does.

Each wrapper names the reference function it restates; the C source carries the
file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)


class OrcCrystals(C.Structure):
    _fields_ = [("n", C.c_int64), ("pos", _f64p), ("dims", _f64p), ("quat", _f64p),
                ("morph", _i32p), ("sens", _f64p), ("developed", _u8p), ("fe", _f64p)]


class OrcBvhOut(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("bmin", _f64p), ("bmax", _f64p), ("left", _u32p),
                ("right", _u32p), ("offset", _u32p), ("count", _u32p), ("indices", _u32p)]


class OrcExposeCfg(C.Structure):
    _fields_ = [("W", C.c_double), ("H", C.c_double), ("D", C.c_double),
                ("mu_a", C.c_double), ("mu_s", C.c_double), ("g", C.c_double),
                ("n_refr", C.c_double), ("r_ahu", C.c_double), ("od_ahu", C.c_double),
                ("mu_abs", C.c_double), ("theta_max", C.c_double),
                ("photons_per_area", C.c_double), ("max_bounces", C.c_int32),
                ("sampler", C.c_int32), ("n_photons", C.c_int64), ("ppp", C.c_int32),
                ("threads", C.c_int32), ("seed", C.c_uint64), ("tw", C.c_int32),
                ("th", C.c_int32), ("pad_", C.c_int32), ("base_thickness", C.c_double)]


class OrcExposeStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "launched", "fresnel_rejected", "zero_weight", "bounces", "nodes_tested",
        "prims_tested", "hits", "abs_events", "abs_total", "chunks", "bounces_live",
        "nodes_live")]


class OrcDevReport(C.Structure):
    _fields_ = [("total", C.c_int64), ("developed", C.c_int64), ("fixed", C.c_int64),
                ("clumps", C.c_int64), ("max_clump", C.c_int64), ("avg_clump", C.c_double)]


class OrcSynthCfg(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "W", "H", "D", "r_min", "mean_radius", "size_sigma", "aspect", "base_sens",
        "sens_sigma", "min_radius", "min_half_thickness")] + [
        ("morph", C.c_int32), ("k_samples", C.c_int32)]


class OrcExpectedCfg(C.Structure):
    _fields_ = [("x", OrcExposeCfg), ("cut", C.c_double), ("fe", C.c_double),
                ("sigma", C.c_double), ("pf", C.c_double), ("w_px", C.c_int32), ("h_px", C.c_int32),
                ("spp", C.c_int32), ("ss", C.c_int32), ("scan_seed", C.c_uint64)]


class OrcFilmPlan(C.Structure):
    _fields_ = [("grain_size_px", C.c_double), ("strength", C.c_double), ("tile_step", C.c_uint32),
                ("pad", C.c_uint32), ("tiles_x", C.c_uint32), ("tiles_y", C.c_uint32)]


class OrcFilmCfg(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("tile_step", C.c_int32), ("pad", C.c_int32),
                ("grain_size_px", C.c_double), ("strength", C.c_float), ("base_nominal", C.c_float),
                ("push", C.c_float), ("scanner_otf", C.c_int32), ("fold", C.c_int32)]


class OrcFilmStats(C.Structure):
    _fields_ = [("high_mean", C.c_float), ("high_std", C.c_float), ("norm_factor", C.c_float),
                ("filter_sigma", C.c_float), ("k_rad", C.c_int32), ("tiles_x", C.c_int32),
                ("tiles_y", C.c_int32), ("pad_", C.c_int32)]


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_crystal_intersect_f64.argtypes = [C.POINTER(OrcCrystals), C.c_int64, _f64p, _f64p, _f64p]
        L.orc_first_hit_f32.argtypes = [C.POINTER(OrcCrystals), C.c_int64, _f32p, _f32p, C.c_int32, _i32p,
                                        _f32p, _f32p]
        L.orc_crystal_intersect_f32.argtypes = [C.POINTER(OrcCrystals), C.c_int64, _f32p, _f32p, _f32p]
        L.orc_bvh_build.argtypes = [C.POINTER(OrcCrystals), C.POINTER(OrcBvhOut)]
        L.orc_bvh_build.restype = C.c_int64
        L.orc_all_hits_f64.argtypes = [C.POINTER(OrcCrystals), C.c_int64, _f64p, _f64p, C.c_int32,
                                       _i32p, _i32p, _f64p, _f64p]
        L.orc_all_hits_f32.argtypes = [C.POINTER(OrcCrystals), C.c_int64, _f32p, _f32p, C.c_int32,
                                       _i32p, _i32p, _f32p, _f32p]
        L.orc_query_radius.argtypes = [C.POINTER(OrcCrystals), _f64p, C.c_double, _i32p, C.c_int64]
        L.orc_query_radius.restype = C.c_int64
        L.orc_expose.argtypes = [C.POINTER(OrcCrystals), _f32p, C.POINTER(OrcExposeCfg), _u32p,
                                 C.POINTER(OrcExposeStats)]
        L.orc_develop.argtypes = [C.POINTER(OrcCrystals), _u32p, C.c_uint32, C.c_double, C.c_double,
                                  _u8p, _f64p, _i64p, C.POINTER(OrcDevReport)]
        L.orc_scan.argtypes = [C.POINTER(OrcCrystals), C.c_double, C.c_double, C.c_double,
                               C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                               C.c_uint64, C.c_int32, _f32p, _i64p]
        L.orc_dneg.argtypes = [_f32p, C.c_int64, _f32p]
        L.orc_otf3.argtypes = [_f32p, C.c_int32, C.c_int32, _f32p]
        L.orc_synthesize.argtypes = [C.POINTER(OrcSynthCfg), C.c_uint64, C.c_int64, _f64p, _f64p,
                                     _f64p, _i32p, _f64p]
        L.orc_synthesize.restype = C.c_int64
        L.orc_darts.argtypes = [C.POINTER(OrcSynthCfg), C.c_uint64, C.c_int64, _f64p, _f64p,
                                _f64p, _i32p, _f64p]
        L.orc_darts.restype = C.c_int64
        L.orc_free_path.argtypes = [C.c_double, C.c_double]
        L.orc_free_path.restype = C.c_double
        L.orc_hg.argtypes = [_f64p, C.c_double, C.c_double, C.c_double, _f64p]
        L.orc_fresnel.argtypes = [C.c_double, C.c_double]
        L.orc_fresnel.restype = C.c_double
        L.orc_antihalation.argtypes = [C.c_double, C.c_double, C.c_uint64, _f64p, _f64p]
        L.orc_expected_dose.argtypes = [C.POINTER(OrcCrystals), _f32p, C.POINTER(OrcExpectedCfg),
                                        _f64p, _i64p]
        L.orc_poisson_tail.argtypes = [_f64p, C.c_int64, C.c_uint32, _f64p, _f64p]
        L.orc_expected_scan.argtypes = [C.POINTER(OrcCrystals), _f64p, _f64p, C.POINTER(OrcExpectedCfg),
                                        _f64p, _f64p, _f64p, _f64p]
        L.orc_touch_graph.argtypes = [C.POINTER(OrcCrystals), C.c_double, C.c_double, _i64p, _i32p,
                                      C.c_int64]
        L.orc_touch_graph.restype = C.c_int64
        L.orc_chord_scaled.argtypes = [C.POINTER(OrcCrystals), C.c_int64, C.c_double, _f64p, _f64p, _f64p]
        L.orc_replay_grad.argtypes = [C.POINTER(OrcCrystals), _f32p, C.POINTER(OrcExpectedCfg),
                                      _f64p, _f64p, _f64p]
        L.orc_hg_score.argtypes = [C.c_double, C.c_double]
        L.orc_hg_score.restype = C.c_double
        L.orc_film_luma.argtypes = [_u8p, C.c_int64, C.c_float, C.c_float, C.c_float, _u8p]
        L.orc_film_plan.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_uint32, C.POINTER(OrcFilmPlan)]
        L.orc_film_optics.argtypes = [_u8p, C.c_int32, C.c_int32, _f32p, _i32p]
        L.orc_film_tile_values.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.orc_film_tile_values.restype = C.c_int64
        L.orc_film_composite.argtypes = [C.POINTER(OrcFilmCfg), _u8p, _f32p, _f32p, _u8p, _f32p, _f32p,
                                         _f32p, _f32p, _f32p, C.POINTER(OrcFilmStats)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


@dataclass
class Crystals:
    """FP64 crystal set (HalideCrystal fields that the path reads)."""
    pos: np.ndarray          # (n,3) f64
    dims: np.ndarray         # (n,3) f64 semi-axes
    quat: np.ndarray         # (n,4) f64 (w,i,j,k)
    morph: np.ndarray        # (n,) i32
    sens: np.ndarray         # (n,) f64
    developed: np.ndarray | None = None   # (n,) u8
    fe: np.ndarray | None = None          # (n,) f64

    def __post_init__(self):
        self.pos = np.ascontiguousarray(self.pos, np.float64).reshape(-1, 3)
        self.dims = np.ascontiguousarray(self.dims, np.float64).reshape(-1, 3)
        self.quat = np.ascontiguousarray(self.quat, np.float64).reshape(-1, 4)
        self.morph = np.ascontiguousarray(self.morph, np.int32).reshape(-1)
        self.sens = np.ascontiguousarray(self.sens, np.float64).reshape(-1)
        if self.developed is not None:
            self.developed = np.ascontiguousarray(self.developed, np.uint8).reshape(-1)
        if self.fe is not None:
            self.fe = np.ascontiguousarray(self.fe, np.float64).reshape(-1)

    @property
    def n(self):
        return self.pos.shape[0]

    def c(self) -> OrcCrystals:
        return OrcCrystals(self.n, _p(self.pos, _f64p), _p(self.dims, _f64p), _p(self.quat, _f64p),
                           _p(self.morph, _i32p), _p(self.sens, _f64p),
                           _p(self.developed, _u8p), _p(self.fe, _f64p))


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(out, _u32p))
    return out


def crystal_intersect(cs: Crystals, i: int, o, d, fp32=False):
    """HalideCrystal::intersect_ray; returns (t0, t1) or None."""
    cc = cs.c()
    if fp32:
        o_, d_ = np.asarray(o, np.float32), np.asarray(d, np.float32)
        t = np.zeros(2, np.float32)
        hit = lib().orc_crystal_intersect_f32(C.byref(cc), i, _p(o_, _f32p), _p(d_, _f32p), _p(t, _f32p))
    else:
        o_, d_ = np.asarray(o, np.float64), np.asarray(d, np.float64)
        t = np.zeros(2, np.float64)
        hit = lib().orc_crystal_intersect_f64(C.byref(cc), i, _p(o_, _f64p), _p(d_, _f64p), _p(t, _f64p))
    return (t[0], t[1]) if hit else None


def bvh_build(cs: Crystals):
    """BVH::build -> dict of node arrays in the reference's pre-order layout."""
    n = cs.n
    cap = max(2 * n, 1)
    bmin = np.zeros((cap, 3)); bmax = np.zeros((cap, 3))
    left = np.zeros(cap, np.uint32); right = np.zeros(cap, np.uint32)
    off = np.zeros(cap, np.uint32); cnt = np.zeros(cap, np.uint32)
    idx = np.zeros(max(n, 1), np.uint32)
    out = OrcBvhOut(0, _p(bmin, _f64p), _p(bmax, _f64p), _p(left, _u32p), _p(right, _u32p),
                    _p(off, _u32p), _p(cnt, _u32p), _p(idx, _u32p))
    cc = cs.c()
    nn = lib().orc_bvh_build(C.byref(cc), C.byref(out))
    return dict(bmin=bmin[:nn], bmax=bmax[:nn], left=left[:nn], right=right[:nn],
                offset=off[:nn], count=cnt[:nn], indices=idx[:n])


def all_hits(cs: Crystals, o, d, max_hits=256, fp32=False):
    """BVH::intersect_ray_all over rays (n,3); returns list of (idx[], t0[], t1[])."""
    ft = np.float32 if fp32 else np.float64
    pt = _f32p if fp32 else _f64p
    o = np.ascontiguousarray(o, ft).reshape(-1, 3)
    d = np.ascontiguousarray(d, ft).reshape(-1, 3)
    nr = o.shape[0]
    nh = np.zeros(nr, np.int32)
    idx = np.zeros((nr, max_hits), np.int32)
    t0 = np.zeros((nr, max_hits), ft)
    t1 = np.zeros((nr, max_hits), ft)
    cc = cs.c()
    fn = lib().orc_all_hits_f32 if fp32 else lib().orc_all_hits_f64
    fn(C.byref(cc), nr, _p(o, pt), _p(d, pt), max_hits, _p(nh, _i32p), _p(idx, _i32p),
       _p(t0, pt), _p(t1, pt))
    if (nh > max_hits).any():
        raise ValueError("max_hits too small")
    return nh, idx, t0, t1


def first_hit(cs: Crystals, o, d, threads=0):
    """First hit per ray (f32 twin of the device traversal): (idx, t0, t1), idx -1 = none;
    the smallest t_enter of intersect_ray_all's hits, ties to the earlier DFS hit."""
    o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
    nr = o.shape[0]
    idx = np.zeros(nr, np.int32)
    t0 = np.zeros(nr, np.float32)
    t1 = np.zeros(nr, np.float32)
    cc = cs.c()
    lib().orc_first_hit_f32(C.byref(cc), nr, _p(o, _f32p), _p(d, _f32p), threads or (os.cpu_count() or 1),
                            _p(idx, _i32p), _p(t0, _f32p), _p(t1, _f32p))
    return idx, t0, t1


def query_radius(cs: Crystals, center, r, cap=4096):
    c = np.asarray(center, np.float64)
    out = np.zeros(cap, np.int32)
    cc = cs.c()
    n = lib().orc_query_radius(C.byref(cc), _p(c, _f64p), r, _p(out, _i32p), cap)
    if n > cap:
        return query_radius(cs, center, r, cap=int(n))
    return out[:n].copy()


def expose(cs: Crystals, exposure_map: np.ndarray | None, *, W, H, D, mu_a, mu_s, g, n_refr,
           r_ahu, od_ahu, mu_abs, theta_max, max_bounces, seed, sampler=1, ppp=0,
           n_photons=0, photons_per_area=0.0, threads=0, base_thickness=0.0):
    """expose_emulsion -> (per-crystal counts u32, stats dict)."""
    if exposure_map is not None:
        em = np.ascontiguousarray(exposure_map, np.float32)
        th, tw = em.shape
    else:
        em, th, tw = None, 1, 1
    cfg = OrcExposeCfg(W, H, D, mu_a, mu_s, g, n_refr, r_ahu, od_ahu, mu_abs, theta_max,
                       photons_per_area, max_bounces, sampler, n_photons, ppp, threads, seed,
                       tw, th, 0, base_thickness)
    counts = np.zeros(max(cs.n, 1), np.uint32)
    st = OrcExposeStats()
    cc = cs.c()
    lib().orc_expose(C.byref(cc), _p(em, _f32p), C.byref(cfg), _p(counts, _u32p), C.byref(st))
    stats = {k: getattr(st, k) for k, _ in OrcExposeStats._fields_}
    return counts[:cs.n], stats


def develop(cs: Crystals, counts, threshold, fe, pf):
    """develop_emulsion -> (developed u8, fe f64, clump_id i64, report dict)."""
    n = cs.n
    counts = np.ascontiguousarray(counts, np.uint32)
    dev = np.zeros(max(n, 1), np.uint8)
    feo = np.zeros(max(n, 1), np.float64)
    cid = np.zeros(max(n, 1), np.int64)
    rep = OrcDevReport()
    cc = cs.c()
    lib().orc_develop(C.byref(cc), _p(counts, _u32p), threshold, fe, pf, _p(dev, _u8p),
                      _p(feo, _f64p), _p(cid, _i64p), C.byref(rep))
    return dev[:n], feo[:n], cid[:n], {k: getattr(rep, k) for k, _ in OrcDevReport._fields_}


def scan(cs: Crystals, W, H, D, w_px, h_px, spp=4, ss=1, sigma=0.25, seed=0, threads=0):
    """scan_emulsion (NegativeTransmittance) -> T (h_px, w_px) f32."""
    T = np.zeros((h_px, w_px), np.float32)
    nodes = C.c_int64(0)
    cc = cs.c()
    lib().orc_scan(C.byref(cc), W, H, D, w_px, h_px, spp, ss, sigma, seed, threads,
                   _p(T, _f32p), C.byref(nodes))
    return T


def dneg(T):
    T = np.ascontiguousarray(T, np.float32)
    out = np.zeros_like(T)
    lib().orc_dneg(_p(T, _f32p), T.size, _p(out, _f32p))
    return out


def otf3(img):
    img = np.ascontiguousarray(img, np.float32)
    h, w = img.shape
    out = np.zeros_like(img)
    lib().orc_otf3(_p(img, _f32p), w, h, _p(out, _f32p))
    return out


# the 17 literal weights of gaussian.wgsl:14-19 (labelled sigma 2.5, radius 8; they sum
# to 1.146, not to 1 as the shader's comment says)
GAUSS17 = np.array([0.002621, 0.007802, 0.019445, 0.040523, 0.070624, 0.103073, 0.125925, 0.134267,
                    0.137440, 0.134267, 0.125925, 0.103073, 0.070624, 0.040523, 0.019445, 0.007802,
                    0.002621], np.float32)


def highpass17(d):
    """delta_d of render_tile_delta_d (gpu.rs:1180-1206): gaussian.wgsl:21-59 as a
    horizontal pass into a temporary and a vertical pass writing orig - blur; taps
    clamped to the edge, each pass an f32 sum in tap order (product, then addition)."""
    d = np.ascontiguousarray(d, np.float32)
    h, w = d.shape
    xs, ys = np.arange(w), np.arange(h)
    tmp = np.zeros_like(d)
    for k in range(-8, 9):
        tmp = tmp + d[:, np.clip(xs + k, 0, w - 1)] * GAUSS17[k + 8]
    acc = np.zeros_like(d)
    for k in range(-8, 9):
        acc = acc + tmp[np.clip(ys + k, 0, h - 1), :] * GAUSS17[k + 8]
    return d - acc


def _synth(fn, *, W, H, D, r_min, mean_radius, size_sigma, morph, aspect=1.0, base_sens=1.2,
               sens_sigma=0.15, min_radius=0.15, min_half_thickness=0.05, k_samples=30, seed=0):
    cfg = OrcSynthCfg(W, H, D, r_min, mean_radius, size_sigma, aspect, base_sens, sens_sigma,
                      min_radius, min_half_thickness, morph, k_samples)
    cap = int(W * H * D / (r_min ** 3) * 1.2) + 64
    while True:
        pos = np.zeros((cap, 3)); dims = np.zeros((cap, 3)); quat = np.zeros((cap, 4))
        mo = np.zeros(cap, np.int32); sens = np.zeros(cap)
        n = fn(C.byref(cfg), seed, cap, _p(pos, _f64p), _p(dims, _f64p), _p(quat, _f64p),
               _p(mo, _i32p), _p(sens, _f64p))
        if n >= 0:
            return Crystals(pos[:n], dims[:n], quat[:n], mo[:n], sens[:n])
        cap = -n


def synthesize(**kw):
    """VirtualEmulsionLayer::synthesize -> Crystals (Bridson, one sequential stream)."""
    return _synth(lib().orc_synthesize, **kw)


def synthesize_darts(**kw):
    """The library's device synthesis (dart throwing) evaluated on the CPU -> Crystals."""
    return _synth(lib().orc_darts, **kw)


def free_path(mu_t, xi):
    """GelatinOptics::sample_free_path for a given uniform."""
    return lib().orc_free_path(mu_t, xi)


def hg(d, g, xi1, xi2):
    """GelatinOptics::sample_henyey_greenstein for given uniforms."""
    d_ = np.asarray(d, np.float64)
    out = np.zeros(3)
    lib().orc_hg(_p(d_, _f64p), g, xi1, xi2, _p(out, _f64p))
    return out


def fresnel(n, cos_i):
    return lib().orc_fresnel(n, cos_i)


def antihalation(r_ahu, od, seed, energy, d):
    e = C.c_double(energy)
    d_ = np.ascontiguousarray(d, np.float64).copy()
    alive = lib().orc_antihalation(r_ahu, od, seed, C.byref(e), _p(d_, _f64p))
    return bool(alive), e.value, d_


# ---------------------------------------------------------------------------------------
# Expected (differentiable) model: FP64 gradient oracle.  No reference counterpart; the
# C source states the model.  The scan mode map and its derivative are evaluated here.

def expected_cfg(exposure_shape, *, W, H, D, mu_a, mu_s, g, n_refr, r_ahu, od_ahu, mu_abs,
                 theta_max, max_bounces, ppp, seed, fe, sigma, w_px, h_px, spp=4, ss=1,
                 scan_seed=0, cut=1e-6, base_thickness=0.0, threads=0, pf=1.25) -> OrcExpectedCfg:
    th, tw = exposure_shape
    x = OrcExposeCfg(W, H, D, mu_a, mu_s, g, n_refr, r_ahu, od_ahu, mu_abs, theta_max, 0.0,
                     max_bounces, 1, 0, ppp, threads, seed, tw, th, 0, base_thickness)
    return OrcExpectedCfg(x, cut, fe, sigma, pf, w_px, h_px, spp, ss, scan_seed)


def poisson_tail(dose, threshold):
    """P[Poisson(dose) >= threshold] and its derivative."""
    d = np.ascontiguousarray(dose, np.float64)
    p = np.zeros_like(d)
    dp = np.zeros_like(d)
    lib().orc_poisson_tail(_p(d, _f64p), d.size, threshold, _p(p, _f64p), _p(dp, _f64p))
    return p, dp


def touch_graph(cs: Crystals, fe, pf):
    """Touch graph of development.rs:132-155 at unboosted expansion: scipy CSR matrix."""
    from scipy.sparse import csr_matrix
    cc = cs.c()
    ptr = np.zeros(cs.n + 1, np.int64)
    cap = max(16 * cs.n, 16)
    while True:
        idx = np.zeros(cap, np.int32)
        k = lib().orc_touch_graph(C.byref(cc), fe, pf, _p(ptr, _i64p), _p(idx, _i32p), cap)
        if k >= 0:
            break
        cap = -k
    return csr_matrix((np.ones(k), idx[:k], ptr), shape=(cs.n, cs.n))


def clump_boost(s):
    """Expected clump boost for expected clump size s (development.rs:164-171: a clump of
    size s >= 2 boosts by 1.15 + 0.18 (min(s, 6) - 1); 1 for s <= 1, linear between) and
    its derivative."""
    s = np.asarray(s, np.float64)
    b = np.where(s <= 1, 1.0, np.where(s <= 2, 1.0 + 0.33 * (s - 1),
                                       1.15 + 0.18 * (np.minimum(s, 6.0) - 1)))
    db = np.where(s <= 1, 0.0, np.where(s <= 2, 0.33, np.where(s < 6, 0.18, 0.0)))
    return b, db


def chord_scaled(cs: Crystals, i, s, o, d):
    out = np.zeros(2)
    cc = cs.c()
    o = np.ascontiguousarray(o, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    hit = lib().orc_chord_scaled(C.byref(cc), i, s, _p(o, _f64p), _p(d, _f64p), _p(out, _f64p))
    return (out[0], out[1]) if hit else None


def mode_map(T, mode):
    T = np.asarray(T, np.float64)
    if mode == 0:
        return T
    if mode == 1:
        return 1.0 - T
    if mode == 2:
        return np.clip(-np.log10(np.maximum(T, 1e-4)) / 3.0, 0.0, 1.0)
    return -np.log10(np.clip(T, 1e-4, 1.0)) + 0.05


def mode_map_grad(T, mode):
    T = np.asarray(T, np.float64)
    ln10 = np.log(10.0)
    if mode == 0:
        return np.ones_like(T)
    if mode == 1:
        return -np.ones_like(T)
    if mode == 2:
        return np.where(T > 1e-3, -1.0 / (3.0 * np.maximum(T, 1e-300) * ln10), 0.0)
    return np.where(T > 1e-4, -1.0 / (np.maximum(T, 1e-300) * ln10), 0.0)


def expected_scan(cs: Crystals, prob, boost, cfg: OrcExpectedCfg, gT=None):
    """T (h, w) for gT None, else (gprob, gboost)."""
    cc = cs.c()
    prob = np.ascontiguousarray(prob, np.float64)
    boost = np.ascontiguousarray(boost, np.float64)
    if gT is None:
        T = np.zeros((cfg.h_px, cfg.w_px), np.float64)
        lib().orc_expected_scan(C.byref(cc), _p(prob, _f64p), _p(boost, _f64p), C.byref(cfg), None,
                                _p(T, _f64p), None, None)
        return T
    gT = np.ascontiguousarray(gT, np.float64)
    gp = np.zeros(max(cs.n, 1))
    gb = np.zeros(max(cs.n, 1))
    lib().orc_expected_scan(C.byref(cc), _p(prob, _f64p), _p(boost, _f64p), C.byref(cfg), _p(gT, _f64p),
                            None, _p(gp, _f64p), _p(gb, _f64p))
    return gp[:cs.n], gb[:cs.n]


def _dev_stage(cs, dose, cfg, threshold, graph):
    prob, dprob = poisson_tail(dose, threshold)
    if graph is None:
        graph = touch_graph(cs, cfg.fe, cfg.pf)
    boost, dboost = clump_boost(1.0 + graph @ prob)
    return prob, dprob, boost, dboost, graph


def expected_forward(cs: Crystals, exposure_map, cfg: OrcExpectedCfg, threshold, mode, graph=None):
    """-> dict(dose, prob, dprob, boost, T, out): the expected model in FP64."""
    em = np.ascontiguousarray(exposure_map, np.float32)
    cc = cs.c()
    dose = np.zeros(max(cs.n, 1), np.float64)
    be = np.zeros(2, np.int64)
    lib().orc_expected_dose(C.byref(cc), _p(em, _f32p), C.byref(cfg), _p(dose, _f64p), _p(be, _i64p))
    dose = dose[:cs.n]
    prob, dprob, boost, dboost, graph = _dev_stage(cs, dose, cfg, threshold, graph)
    T = expected_scan(cs, prob, boost, cfg)
    return {"dose": dose, "prob": prob, "dprob": dprob, "boost": boost, "T": T,
            "out": mode_map(T, mode), "bounces": int(be[0]), "events": int(be[1]), "graph": graph}


def expected_vjp(cs: Crystals, exposure_map, cfg: OrcExpectedCfg, threshold, mode, grad_out,
                 dose=None, graph=None):
    """-> dict(grad_exposure, grads [mu_abs, g, R], grad_dose, grad_prob) in FP64, for
    dL/dout = grad_out; dose from a forward (recomputed when None)."""
    em = np.ascontiguousarray(exposure_map, np.float32)
    cc = cs.c()
    if dose is None:
        dose = expected_forward(cs, em, cfg, threshold, mode, graph)["dose"]
    dose = np.ascontiguousarray(dose, np.float64)
    prob, dprob, boost, dboost, graph = _dev_stage(cs, dose, cfg, threshold, graph)
    T = expected_scan(cs, prob, boost, cfg)
    gT = np.asarray(grad_out, np.float64) * mode_map_grad(T, mode)
    gprob, gboost = expected_scan(cs, prob, boost, cfg, gT)
    gprob = gprob + graph.T @ (gboost * dboost)
    gdose = np.ascontiguousarray(gprob * dprob)
    gmap = np.zeros(em.shape, np.float64)
    grads = np.zeros(3, np.float64)
    lib().orc_replay_grad(C.byref(cc), _p(em, _f32p), C.byref(cfg), _p(gdose, _f64p),
                          _p(gmap, _f64p), _p(grads, _f64p))
    return {"grad_exposure": gmap, "grads": grads, "grad_dose": gdose, "grad_prob": gprob,
            "grad_boost": gboost}


def hg_score(g, mu):
    return lib().orc_hg_score(g, mu)


# ------------------------------------------------------------- film compositing -----
# Per-preset constants of render_film_black_and_white_tiled (image_pipeline.rs:264-283,
# :726-731).
FILM_LUMA_WEIGHTS = {
    "tri-x": (0.42, 0.46, 0.12), "trix": (0.42, 0.46, 0.12), "tri-x-400": (0.42, 0.46, 0.12),
    "tri-x-1600": (0.42, 0.46, 0.12), "trix-1600": (0.42, 0.46, 0.12), "caffenol": (0.42, 0.46, 0.12),
    "hp5": (0.36, 0.48, 0.16), "hp5-plus": (0.36, 0.48, 0.16),
    "t-max": (0.33, 0.52, 0.15), "tmax": (0.33, 0.52, 0.15),
}


def film_luma_weights(preset: str):
    return FILM_LUMA_WEIGHTS.get(preset.lower(), (0.38, 0.48, 0.14))


def film_base_nominal(preset: str) -> float:
    p = preset.lower()
    if p in ("t-max", "tmax"):
        return 0.024
    if p in ("hp5", "hp5-plus"):
        return 0.032
    if p in ("tri-x-1600", "trix-1600", "caffenol"):
        return 0.042
    return 0.038


def film_effective_push(preset: str, push: float) -> float:
    p = preset.lower()
    return max(push, 2.0) if ("1600" in p or p == "caffenol") else max(push, 0.0)


def film_luma(rgb, preset: str):
    """convert_to_panchromatic_luma of an (h, w, 3) u8 image."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    out = np.zeros(rgb.shape[:2], np.uint8)
    rw, gw, bw = film_luma_weights(preset)
    lib().orc_film_luma(_p(rgb, _u8p), out.size, rw, gw, bw, _p(out, _u8p))
    return out


def film_plan(w, h, grain_size_px=2.2, grain_strength=1.0, scale=1.0, auto_scale=False,
              reference_res=6000, tile_size=320) -> dict:
    p = OrcFilmPlan()
    if lib().orc_film_plan(grain_size_px, grain_strength, scale, int(auto_scale), reference_res,
                           tile_size, w, h, C.byref(p)) != 0:
        raise ValueError("tile_size below 48")
    return {k: getattr(p, k) for k, _ in OrcFilmPlan._fields_}


def film_optics(src):
    """apply_camera_and_gelatin_optics: (h_effective f32, (top, bottom, left, right))."""
    src = np.ascontiguousarray(src, np.uint8)
    h, w = src.shape
    out = np.zeros((h, w), np.float32)
    m = np.zeros(4, np.int32)
    lib().orc_film_optics(_p(src, _u8p), w, h, _p(out, _f32p), _p(m, _i32p))
    return out, tuple(int(v) for v in m)


def film_tile_spans(w, h, step, pad):
    """(x_start, y_start, tw, th) per tile in tile order (image_pipeline.rs:397-421)."""
    out = []
    for ty in range(-(-h // step)):
        for tx in range(-(-w // step)):
            x0, y0 = tx * step, ty * step
            xs, ys = max(x0 - pad, 0), max(y0 - pad, 0)
            out.append((xs, ys, min(x0 + step + pad, w) - xs, min(y0 + step + pad, h) - ys))
    return out


def film_composite(src, h_eff, tiles, *, tile_step, pad, grain_size_px, strength=1.0,
                   preset="tri-x", push=0.0, scanner_otf=True, fold=0, intermediates=False):
    """Compositing of the rendered tiles (image_pipeline.rs:593-891).  tiles: concatenated
    per-tile D_neg in tile order.  Returns (u8 image, stats dict[, intermediates dict])."""
    src = np.ascontiguousarray(src, np.uint8)
    h, w = src.shape
    h_eff = np.ascontiguousarray(h_eff, np.float32)
    tiles = np.ascontiguousarray(tiles, np.float32).reshape(-1)
    need = lib().orc_film_tile_values(w, h, tile_step, pad)
    if tiles.size != need:
        raise ValueError(f"tiles hold {tiles.size} values, layout needs {need}")
    c = OrcFilmCfg(w, h, tile_step, pad, grain_size_px, strength, film_base_nominal(preset),
                   film_effective_push(preset, push), int(bool(scanner_otf)), fold)
    out = np.zeros((h, w), np.uint8)
    bufs = {k: np.zeros((h, w), np.float32) for k in ("d_norm", "d_low", "grain", "scanned", "tone")} \
        if intermediates else {}
    st = OrcFilmStats()
    r = lib().orc_film_composite(C.byref(c), _p(src, _u8p), _p(h_eff, _f32p), _p(tiles, _f32p),
                                 _p(out, _u8p), *[_p(bufs.get(k), _f32p) for k in
                                                  ("d_norm", "d_low", "grain", "scanned", "tone")],
                                 C.byref(st))
    if r != 0:
        raise ValueError("bad compositing layout")
    stats = {k: getattr(st, k) for k, _ in OrcFilmStats._fields_ if k != "pad_"}
    return (out, stats, bufs) if intermediates else (out, stats)
