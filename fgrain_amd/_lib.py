"""ctypes binding of libfgrain_hip.so (include/fgrain_hip.h).

The library is the product: there is no Python or CPU fallback.  Loading fails with
an ImportError naming the missing file when the HIP library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfgrain_hip.so")

FG_OK = 0
STATUS_NAMES = {
    0: "Ok", 1: "AdapterNotFound", 2: "DeviceRequestFailed", 3: "ChannelError",
    4: "BufferMapError", 5: "BufferRangeError", 6: "PipelineError", 7: "InvalidArgument",
}

_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)


class Optics(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "refractive_index", "absorption_coeff", "scattering_coeff", "anisotropy_g",
        "antihalation_reflectance", "antihalation_density")]


class EmulsionConfigC(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "width", "height", "depth", "r_min", "mean_crystal_radius", "size_variance",
        "tabular_aspect_ratio", "base_sensitivity", "sensitivity_sigma", "min_radius",
        "min_half_thickness")] + [("morphology", C.c_uint32), ("k_samples", C.c_uint32),
                                  ("optics", Optics)]


class ExposureConfigC(C.Structure):
    _fields_ = [("photons_per_sq_unit", C.c_double), ("illumination_angle_rad", C.c_double),
                ("crystal_absorption_coeff", C.c_double), ("max_bounces", C.c_uint32),
                ("photons_per_pixel", C.c_uint32), ("base_thickness", C.c_double)]


class ExposeStatsC(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "launched", "bounces", "nodes_visited", "prims_tested", "abs_events", "abs_total")] + [
        ("kernel_ms", C.c_float), ("pad_", C.c_float), ("atomics", C.c_uint64), ("packets_run", C.c_uint64)]


class DeviceTileC(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("records", C.c_void_p), ("counts", C.c_void_p),
                ("exposure", C.c_void_p), ("n_nodes", C.c_uint32), ("n_crystals", C.c_uint32),
                ("tw", C.c_uint32), ("th", C.c_uint32)]


class DevelopmentConfigC(C.Structure):
    _fields_ = [("threshold_photons", C.c_uint32), ("pad_", C.c_uint32),
                ("filament_expansion", C.c_double), ("clumping_proximity_factor", C.c_double),
                ("developer_contrast", C.c_double)]


class DevelopmentReportC(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "total_crystals", "developed_crystals", "fixed_crystals", "total_clumps",
        "max_clump_size")] + [("avg_clump_size", C.c_double)]


class ScanConfigC(C.Structure):
    _fields_ = [("width_px", C.c_uint32), ("height_px", C.c_uint32),
                ("samples_per_pixel", C.c_uint32), ("mode", C.c_uint32),
                ("silver_opacity", C.c_double), ("supersample", C.c_uint32), ("otf", C.c_uint32)]


class TuningC(C.Structure):
    _fields_ = [("transport_kernel", C.c_uint32), ("transport_chunk", C.c_uint32),
                ("transport_wait_pct", C.c_uint32), ("scan_rays_per_thread", C.c_uint32),
                ("transport_waves", C.c_uint32), ("transport_prereduce", C.c_uint32)]


class TileParamsC(C.Structure):
    _fields_ = [("film_w", C.c_double), ("film_h", C.c_double), ("film_d", C.c_double),
                ("width_px", C.c_uint32), ("height_px", C.c_uint32),
                ("samples_per_pixel", C.c_uint32), ("morphology", C.c_uint32),
                ("silver_opacity", C.c_double), ("seed", C.c_uint64),
                ("r_min", C.c_double), ("mean_crystal_radius", C.c_double),
                ("size_variance", C.c_double), ("tabular_aspect_ratio", C.c_double),
                ("base_sensitivity", C.c_double), ("exposure", C.c_double),
                ("threshold_photons", C.c_uint32), ("max_bounces", C.c_uint32),
                ("filament_expansion", C.c_double), ("clumping_proximity_factor", C.c_double),
                ("optics", Optics), ("photons_per_sq_unit", C.c_double),
                ("illumination_angle_rad", C.c_double), ("crystal_absorption_coeff", C.c_double),
                ("depth", C.c_double)]


class TileStatsC(C.Structure):
    _fields_ = [("crystals", C.c_uint64), ("developed", C.c_uint64), ("clumps", C.c_uint64),
                ("photons", C.c_uint64), ("ms_transport", C.c_float), ("ms_develop", C.c_float),
                ("ms_scan", C.c_float), ("ms_total", C.c_float)]


class FrameDescC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("tile", C.c_uint32),
                ("threads", C.c_uint32), ("pixel_size", C.c_double),
                ("emulsion", EmulsionConfigC), ("exposure", ExposureConfigC),
                ("development", DevelopmentConfigC), ("scan", ScanConfigC),
                ("seed", C.c_uint64), ("exposure_map", _f32p), ("density_mult", _f32p),
                ("synthesis", C.c_uint32), ("overlap", C.c_uint32)]


class FrameStatsC(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "tiles", "crystals", "developed", "clumps", "photons", "rays", "pixels", "bounces",
        "nodes_visited", "prims_tested", "abs_events")] + [
        (n, C.c_float) for n in ("ms_total", "ms_transport", "ms_develop", "ms_scan",
                                 "ms_kernel_transport", "ms_kernel_scan")] + [
        ("graph_used", C.c_uint32), ("pad_", C.c_uint32), ("atomics", C.c_uint64), ("packets_run", C.c_uint64)]


class ExpectedModelC(C.Structure):
    _fields_ = [("exposure", ExposureConfigC), ("optics", Optics),
                ("development", DevelopmentConfigC), ("scan", ScanConfigC),
                ("expose_seed", C.c_uint64), ("scan_seed", C.c_uint64),
                ("weight_cutoff", C.c_double)]


class ParamGradsC(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "crystal_absorption_coeff", "anisotropy_g", "antihalation_reflectance", "reserved_")]


class AdjointStatsC(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("packets", "bounces", "events")] + [
        (n, C.c_float) for n in ("ms_replay", "ms_scan", "ms_total", "pad_")] + [("atomics", C.c_uint64)]


class FilmOptionsC(C.Structure):
    _fields_ = [("preset", C.c_char * 32), ("grain_size_px", C.c_double), ("grain_strength", C.c_double),
                ("exposure", C.c_double), ("scale", C.c_double), ("push", C.c_double), ("seed", C.c_uint64),
                ("tile_size", C.c_uint32), ("reference_res", C.c_uint32), ("auto_scale", C.c_uint32),
                ("scanner_otf", C.c_uint32), ("synthesis", C.c_uint32), ("frame_mode", C.c_uint32),
                ("threads", C.c_uint32), ("pad_", C.c_uint32)]


class FilmPlanC(C.Structure):
    _fields_ = [("grain_size_px", C.c_double), ("strength", C.c_double), ("push", C.c_double),
                ("pixel_size_um", C.c_double), ("tile_step", C.c_uint32), ("pad", C.c_uint32),
                ("tiles_x", C.c_uint32), ("tiles_y", C.c_uint32), ("filter_sigma", C.c_float),
                ("k_rad", C.c_uint32), ("base_nominal", C.c_float), ("pad_", C.c_uint32)]


class FilmReportC(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("in_w", "in_h", "out_w", "out_h")] + [
        (n, C.c_uint64) for n in ("total_tiles", "total_crystals", "developed_crystals", "total_clumps",
                                  "photons", "rays")] + [("elapsed_seconds", C.c_double)] + [
        (n, C.c_float) for n in ("high_mean", "high_std", "norm_factor", "pad_", "ms_luma_optics",
                                 "ms_synthesis", "ms_render", "ms_composite", "ms_kernel_optics",
                                 "ms_kernel_composite", "ms_kernel_transport", "ms_kernel_scan")]


class CompositeStatsC(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("high_mean", "high_std", "norm_factor", "filter_sigma")] + [
        (n, C.c_uint32) for n in ("k_rad", "tiles_x", "tiles_y", "pad_")] + [
        ("ms_kernels", C.c_float), ("pad2_", C.c_float)]


# every entry point of include/fgrain_hip.h: name -> (restype, argtypes)
_vp = C.c_void_p
_st = C.c_int
SIGNATURES = {
    "fg_abi_version": (C.c_uint32, []),
    "fg_last_error": (C.c_char_p, []),
    "fg_engine_create": (_st, [C.POINTER(_vp)]),
    "fg_engine_destroy": (None, [_vp]),
    "fg_engine_adapter_name": (C.c_char_p, [_vp]),
    "fg_get_engine": (_st, [C.POINTER(_vp)]),
    "fg_engine_set_tuning": (_st, [_vp, C.POINTER(TuningC)]),
    "fg_engine_get_tuning": (_st, [_vp, C.POINTER(TuningC)]),
    "fg_calibrate_valu": (_st, [_vp, C.c_uint32, C.c_uint32, _f32p, C.POINTER(C.c_uint64)]),
    "fg_emulsion_synthesize": (_st, [C.POINTER(EmulsionConfigC), C.c_uint64, C.POINTER(_vp)]),
    "fg_emulsion_synthesize_device": (_st, [_vp, C.POINTER(EmulsionConfigC), C.c_uint64, C.POINTER(_vp)]),
    "fg_emulsion_from_crystals": (_st, [C.POINTER(EmulsionConfigC), C.c_int64, _f64p, _f64p,
                                        _f64p, _i32p, _f64p, C.POINTER(_vp)]),
    "fg_emulsion_destroy": (None, [_vp]),
    "fg_emulsion_count": (C.c_int64, [_vp]),
    "fg_emulsion_get_crystals": (_st, [_vp, _f64p, _f64p, _f64p, _i32p, _f64p]),
    "fg_emulsion_get_bvh": (_st, [_vp, _i64p, _f64p, _f64p, _u32p, _u32p, _u32p, _u32p, _u32p]),
    "fg_emulsion_get_device_layout": (_st, [_vp, _vp, _vp]),
    "fg_emulsion_get_state": (_st, [_vp, _u32p, _u8p, _f64p, _i64p]),
    "fg_emulsion_set_state": (_st, [_vp, _u32p, _u8p, _f64p]),
    "fg_expose": (_st, [_vp, _vp, _f32p, C.c_uint32, C.c_uint32, C.POINTER(ExposureConfigC),
                        C.c_uint64, C.POINTER(ExposeStatsC)]),
    "fg_expose_noalloc": (_st, [_vp, _vp, C.POINTER(DeviceTileC), C.POINTER(ExposureConfigC), C.c_uint64,
                                _vp, C.POINTER(ExposeStatsC)]),
    "fg_develop": (_st, [_vp, _vp, C.POINTER(DevelopmentConfigC), C.POINTER(DevelopmentReportC)]),
    "fg_scan": (_st, [_vp, _vp, C.POINTER(ScanConfigC), C.c_uint64, _f32p]),
    "fg_traverse_all_hits": (_st, [_vp, _vp, C.c_int64, _f32p, _f32p, C.c_int32, _i32p, _i32p,
                                   _f32p, _f32p]),
    "fg_traverse_first_hit": (_st, [_vp, _vp, C.c_int64, _f32p, _f32p, _i32p, _f32p, _f32p]),
    "fg_tile_params_init": (_st, [C.POINTER(TileParamsC), C.c_char_p, C.c_double, C.c_uint32,
                                  C.c_uint32, C.c_double, C.c_double, C.c_uint64]),
    "fg_render_tile_all_gpu": (_st, [_vp, _f32p, C.POINTER(TileParamsC), _f32p,
                                     C.POINTER(TileStatsC)]),
    "fg_render_tile_delta_d": (_st, [_vp, C.c_int64, _f64p, _f64p, _f64p, _i32p, _f64p, C.POINTER(TileParamsC),
                                     _f32p, _f32p]),
    "fg_frame_create": (_st, [_vp, C.POINTER(FrameDescC), C.POINTER(_vp)]),
    "fg_frame_destroy": (None, [_vp]),
    "fg_frame_render": (_st, [_vp, _vp, C.c_uint32, C.POINTER(FrameStatsC)]),
    "fg_frame_read": (_st, [_vp, _vp, _f32p]),
    "fg_frame_read_tiles": (_st, [_vp, _vp, C.c_int64, _f32p, _i64p]),
    "fg_frame_read_counts": (_st, [_vp, _vp, C.c_int64, _u32p]),
    "fg_frame_read_development": (_st, [_vp, _vp, C.c_int64, _u8p, _f64p, _i64p]),
    "fg_frame_get_tile": (_st, [_vp, _vp, C.c_int64, _i64p, _i64p, _f64p, _f64p, _f64p, _i32p, _f64p, _vp, _vp]),
    "fg_frame_crystal_count": (C.c_int64, [_vp]),
    "fg_expected_forward": (_st, [_vp, _vp, C.POINTER(ExpectedModelC), _f32p, C.c_uint32,
                                  C.c_uint32, _f32p, _f32p, _f32p, C.POINTER(AdjointStatsC)]),
    "fg_vjp": (_st, [_vp, _vp, C.POINTER(ExpectedModelC), _f32p, C.c_uint32, C.c_uint32,
                     _f32p, _f32p, _f32p, C.POINTER(ParamGradsC), _f32p,
                     C.POINTER(AdjointStatsC)]),
    "fg_film_options_default": (_st, [C.POINTER(FilmOptionsC)]),
    "fg_film_plan_for": (_st, [C.POINTER(FilmOptionsC), C.c_uint32, C.c_uint32, C.POINTER(FilmPlanC)]),
    "fg_render_film_bw": (_st, [_vp, _u8p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(FilmOptionsC),
                                _u8p, C.POINTER(FilmReportC)]),
    "fg_film_luma": (_st, [_vp, _u8p, C.c_uint32, C.c_uint32, C.c_char_p, _u8p, _f32p]),
    "fg_film_optics": (_st, [_vp, _u8p, C.c_uint32, C.c_uint32, _f32p, _f32p]),
    "fg_film_composite": (_st, [_vp, C.POINTER(FilmOptionsC), C.c_uint32, C.c_uint32, _u8p, _f32p, _f32p,
                                C.c_int64, _u8p, _f32p, _f32p, _f32p, C.POINTER(CompositeStatsC)]),
}

_lib = None


def lib():
    """The loaded library; raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"fgrain_amd: HIP library missing at {LIB_PATH}; "
                              "run __graft_entry__.build() (hipcc, gfx950) first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class GpuError(RuntimeError):
    """GpuError (src/error.rs:9-33): status code + payload string."""

    def __init__(self, status: int, msg: str):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")


def check(status: int):
    if status != FG_OK:
        raise GpuError(status, lib().fg_last_error().decode(errors="replace"))


def ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)
