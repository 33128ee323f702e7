/*
 * fgrain_hip.h -- C ABI of the MI355X (gfx950, HIP) photon-transport + scan engine.
 *
 * Drop-in for fgrain's GPU boundary (GuyNguyen/fgrain src/gpu.rs) and for the library
 * stages the `--gpu` path runs (emulsion synthesis -> transport -> development ->
 * transmittance scan, src/lib.rs:89-111).  Plain C types only: status codes are
 * returned, outputs are caller-allocated, nothing throws across the ABI.  All lengths
 * are in one caller-chosen unit (micrometres in the reference's image pipeline,
 * pixels in the synthetic workloads); coefficients are per that unit.
 *
 * Threading: every entry point may be called from any thread; calls on one engine are
 * serialised by an internal mutex (the reference's dispatch_lock, gpu.rs:169, 503).
 */
#ifndef FGRAIN_HIP_H
#define FGRAIN_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FG_ABI_VERSION 1u

/* GpuError variants (src/error.rs:9-33) plus argument validation. */
typedef enum fg_status {
    FG_OK = 0,
    FG_ERR_ADAPTER_NOT_FOUND = 1, /* GpuError::AdapterNotFound */
    FG_ERR_DEVICE_REQUEST = 2,    /* GpuError::DeviceRequestFailed */
    FG_ERR_CHANNEL = 3,           /* GpuError::ChannelError */
    FG_ERR_BUFFER_MAP = 4,        /* GpuError::BufferMapError */
    FG_ERR_BUFFER_RANGE = 5,      /* GpuError::BufferRangeError */
    FG_ERR_PIPELINE = 6,          /* GpuError::PipelineError */
    FG_ERR_INVALID_ARG = 7
} fg_status;

typedef struct fg_engine fg_engine;     /* opaque, thread-safe */
typedef struct fg_emulsion fg_emulsion; /* VirtualEmulsionLayer (emulsion.rs:141-150) */
typedef struct fg_frame fg_frame;       /* device-resident batch of tiles */

/* GelatinOptics (src/optics.rs:41-60) */
typedef struct fg_optics {
    double refractive_index;
    double absorption_coeff;
    double scattering_coeff;
    double anisotropy_g;
    double antihalation_reflectance;
    double antihalation_density;
} fg_optics;

/* EmulsionConfig (src/emulsion.rs:14-38).  min_radius and min_half_thickness are the
 * reference's fixed clamps 0.15 and 0.05 (emulsion.rs:169, 187), in the length unit. */
typedef struct fg_emulsion_config {
    double width, height, depth;
    double r_min;
    double mean_crystal_radius;
    double size_variance;
    double tabular_aspect_ratio;
    double base_sensitivity;
    double sensitivity_sigma;   /* 0.15 in the reference (emulsion.rs:164) */
    double min_radius;
    double min_half_thickness;
    uint32_t morphology;        /* 0 cubic, 1 tabular, 2 octahedral */
    uint32_t k_samples;         /* Bridson candidates, 30 (emulsion.rs:159) */
    fg_optics optics;
} fg_emulsion_config;

/* ExposureConfig (src/transport.rs:15-39).  photons_per_pixel > 0 selects the
 * stratified sampler (that many packets per exposure-map pixel, positions jittered in
 * the pixel); 0 selects uniform positions over the patch with
 * round(width*height*photons_per_sq_unit) packets. */
typedef struct fg_exposure_config {
    double photons_per_sq_unit;
    double illumination_angle_rad;
    double crystal_absorption_coeff;
    uint32_t max_bounces;
    uint32_t photons_per_pixel;
    double base_thickness;  /* film base below the emulsion; 0 = reflect in place as the
                               reference (optics.rs:160-196); > 0: a reflected packet
                               re-enters displaced by base * (tan in + tan out) */
} fg_exposure_config;

typedef struct fg_expose_stats {
    uint64_t launched;      /* packets launched (Fresnel-rejected included) */
    uint64_t bounces;
    uint64_t nodes_visited;
    uint64_t prims_tested;
    uint64_t abs_events;    /* hits that deposited k > 0 latent photons */
    uint64_t abs_total;     /* sum of deposited latent photons */
    float kernel_ms;        /* transport kernel time (HIP events) */
    float pad_;
    uint64_t atomics;       /* global atomic additions issued for the deposits after the
                               wave pre-reduction (abs_events without it) */
    uint64_t packets_run;   /* packets the persistent kernel started (= launched; 0 for
                               the one-thread-per-packet kernel) */
} fg_expose_stats;

/* DevelopmentConfig (src/development.rs:12-38) */
typedef struct fg_development_config {
    uint32_t threshold_photons;
    uint32_t pad_;
    double filament_expansion;
    double clumping_proximity_factor;
    double developer_contrast;
} fg_development_config;

/* DevelopmentReport (src/development.rs:42-55) */
typedef struct fg_development_report {
    uint64_t total_crystals;
    uint64_t developed_crystals;
    uint64_t fixed_crystals;
    uint64_t total_clumps;
    uint64_t max_clump_size;
    double avg_clump_size;
} fg_development_report;

#define FG_SCAN_NEGATIVE 0u  /* ScanMode::NegativeTransmittance */
#define FG_SCAN_POSITIVE 1u  /* ScanMode::PositivePrint */
#define FG_SCAN_DENSITY 2u   /* ScanMode::OpticalDensity */
#define FG_SCAN_DNEG 3u      /* D_neg = -log10(clamp(T,1e-4,1)) + 0.05 (image_pipeline.rs:1006-1011) */

/* ScanConfig (src/transmittance.rs:32-60) + aperture supersampling and scanner OTF.
 * supersample = s splits each output pixel into s x s apertures with
 * round(sqrt(spp))^2 stratified rays each; otf = 1 applies the separable
 * [0.035, 0.93, 0.035] filter with clamp-to-edge (image_pipeline.rs:822-855). */
typedef struct fg_scan_config {
    uint32_t width_px, height_px;
    uint32_t samples_per_pixel;
    uint32_t mode;
    double silver_opacity;
    uint32_t supersample;
    uint32_t otf;
} fg_scan_config;

/* GpuTileParams (src/gpu.rs:18-37), plus the transport and optics values the
 * reference's CPU tile path takes from its preset tables (image_pipeline.rs:924-976). */
typedef struct fg_tile_params {
    double film_w, film_h, film_d;
    uint32_t width_px, height_px;
    uint32_t samples_per_pixel;
    uint32_t morphology;
    double silver_opacity;
    uint64_t seed;
    double r_min, mean_crystal_radius, size_variance, tabular_aspect_ratio, base_sensitivity;
    double exposure;
    uint32_t threshold_photons;
    uint32_t max_bounces;
    double filament_expansion, clumping_proximity_factor;
    fg_optics optics;
    double photons_per_sq_unit;     /* 6 * max(exposure, 0.05) in the reference */
    double illumination_angle_rad;  /* 0.12 */
    double crystal_absorption_coeff;/* 1.3 */
    double depth;                   /* emulsion depth used for synthesis (= film_d) */
} fg_tile_params;

typedef struct fg_tile_stats {
    uint64_t crystals, developed, clumps;
    uint64_t photons;
    float ms_transport, ms_develop, ms_scan, ms_total;
} fg_tile_stats;

/* ---------------------------------------------------------------- engine ------- */
/* GpuEngine::new (gpu.rs:174).  No HIP device -> FG_ERR_ADAPTER_NOT_FOUND. */
fg_status fg_engine_create(fg_engine **out);
void fg_engine_destroy(fg_engine *engine);
/* GpuEngine::adapter_name (gpu.rs:484-486) */
const char *fg_engine_adapter_name(const fg_engine *engine);
/* get_gpu_engine (gpu.rs:9-14): process-wide engine created on first use; a failed
 * creation is remembered and reported on every later call. */
fg_status fg_get_engine(fg_engine **out);
/* Message of the last failed call on this thread (GpuError's payload string). */
const char *fg_last_error(void);
uint32_t fg_abi_version(void);

/* Launch parameters chosen by the autotuner (fgrain_amd/autotune.py); zero fields keep
 * the defaults.  They change speed only: transport settings give bit-identical results;
 * the scan shape changes only the summation order of a pixel's rays (float rounding). */
typedef struct fg_tuning {
    uint32_t transport_kernel;   /* 0 persistent lanes (default), 1 one thread per packet;
                                    per-tile calls and the replay (batched frames: persistent) */
    uint32_t transport_chunk;    /* packets per work fetch of a wave, >= 64 (default 64) */
    uint32_t transport_wait_pct; /* BVH steps continue until this % of lanes hold a leaf (80) */
    uint32_t scan_rays_per_thread; /* scan rays per lane: 0 or 1 one lane per ray (default);
                                      r > 1 gives a pixel rays / r adjacent lanes (<= 64) */
    uint32_t transport_waves;    /* waves per SIMD of the batched frame transport: 6 (80
                                    VGPRs) or 7 (72 VGPRs, more spills); 0 means 6 */
    uint32_t transport_prereduce;/* 1: latent-photon deposits of the persistent transport
                                    are pre-reduced across each wave before the global
                                    atomics (39 % fewer atomics on the Tri-X frame, 4-6 %
                                    more transport time); 0 (default): one atomic per
                                    absorbing hit.  Same counts.  The dose replay of the
                                    adjoint always pre-reduces its f64 deposits. */
} fg_tuning;
fg_status fg_engine_set_tuning(fg_engine *engine, const fg_tuning *tuning);
fg_status fg_engine_get_tuning(const fg_engine *engine, fg_tuning *tuning);

/* Measurement support (no reference counterpart): runs blocks x 256 threads of a kernel
 * that issues exactly 64 * iters + 4 VALU instructions per wave (*valu_per_wave), so a
 * profiler's VALU instruction counter can be scaled and the device's issue rate read
 * from *ms_kernel (HIP events). */
fg_status fg_calibrate_valu(fg_engine *engine, uint32_t blocks, uint32_t iters, float *ms_kernel,
                            uint64_t *valu_per_wave);

/* ---------------------------------------------------------------- emulsion ----- */
/* VirtualEmulsionLayer::synthesize (emulsion.rs:155-220): Bridson 3D + per-crystal
 * draws + BVH::build.  Host work; deterministic in seed. */
fg_status fg_emulsion_synthesize(const fg_emulsion_config *cfg, uint64_t seed,
                                 fg_emulsion **out);
/* An emulsion from given crystals (dims = semi-axes, quat = (w,i,j,k)). */
/* Device synthesis (no reference counterpart: replaces the serial Bridson sampler of
 * poisson.rs:59-129 when speed matters).  Poisson-disk centres by parallel dart throwing
 * on the GPU: a grid of cell r_min / sqrt(3), 27 phases of cells at least two cells
 * apart, k_samples rounds of one candidate per empty cell; minimum distance r_min like
 * the reference, density somewhat higher than Bridson's.  Crystal attributes follow
 * emulsion.rs:160-207 with one Philox stream per crystal; the BVH is the reference's. */
fg_status fg_emulsion_synthesize_device(fg_engine *engine, const fg_emulsion_config *cfg,
                                        uint64_t seed, fg_emulsion **out);
fg_status fg_emulsion_from_crystals(const fg_emulsion_config *cfg, int64_t n,
                                    const double *pos, const double *dims,
                                    const double *quat, const int32_t *morph,
                                    const double *sens, fg_emulsion **out);
void fg_emulsion_destroy(fg_emulsion *em);
int64_t fg_emulsion_count(const fg_emulsion *em);
fg_status fg_emulsion_get_crystals(const fg_emulsion *em, double *pos, double *dims,
                                   double *quat, int32_t *morph, double *sens);
/* BVH::build layout (bvh.rs:129-151): node boxes (f64), children, leaf ranges,
 * indices.  Pass NULL pointers to query n_nodes only. */
fg_status fg_emulsion_get_bvh(const fg_emulsion *em, int64_t *n_nodes, double *bmin,
                              double *bmax, uint32_t *left, uint32_t *right,
                              uint32_t *offset, uint32_t *count, uint32_t *indices);
/* Threaded 32-byte device layout (n_nodes * 32 bytes) and 48-byte leaf-order records. */
fg_status fg_emulsion_get_device_layout(const fg_emulsion *em, void *nodes, void *records);
/* Latent counts (HalideCrystal::absorbed_photons), developed flags, filament expansion
 * and clump ids (-1 = none), per crystal in the caller's crystal order. */
fg_status fg_emulsion_get_state(const fg_emulsion *em, uint32_t *counts, uint8_t *developed,
                                double *fe, int64_t *clump_id);
fg_status fg_emulsion_set_state(fg_emulsion *em, const uint32_t *counts,
                                const uint8_t *developed, const double *fe);

/* ---------------------------------------------------------------- stages ------- */
/* expose_emulsion (transport.rs:48-188) on the GPU: adds latent photons to the
 * emulsion's counts.  exposure_map is a tw x th row-major f32 map looked up by the
 * tile path's nearest-pixel rule (image_pipeline.rs:949-961); NULL = uniform 1. */
fg_status fg_expose(fg_engine *engine, fg_emulsion *em, const float *exposure_map,
                    uint32_t tw, uint32_t th, const fg_exposure_config *cfg, uint64_t seed,
                    fg_expose_stats *stats);
/* develop_emulsion (development.rs:59-188): integer threshold (exact), clumping as
 * connected components, clump boost, from the emulsion's latent counts.  With an engine
 * it runs on the device (the batched frame's development kernels); engine NULL runs the
 * same rules on the host. */
fg_status fg_develop(fg_engine *engine, fg_emulsion *em, const fg_development_config *cfg,
                     fg_development_report *report);
/* scan_emulsion (transmittance.rs:65-171) over the developed crystals; out has
 * width_px * height_px values in the configured mode. */
fg_status fg_scan(fg_engine *engine, const fg_emulsion *em, const fg_scan_config *cfg,
                  uint64_t seed, float *out);
/* BVH::intersect_ray_all (bvh.rs:300-347) on the GPU for n_rays float rays: hits in
 * the reference's DFS callback order (crystal index, t_enter, t_exit). */
fg_status fg_traverse_all_hits(fg_engine *engine, const fg_emulsion *em, int64_t n_rays,
                               const float *origins, const float *dirs, int32_t max_hits,
                               int32_t *n_hits, int32_t *idx, float *t0, float *t1);
/* First hit = smallest t_enter, ties to the earlier hit in DFS order; idx -1 = none. */
fg_status fg_traverse_first_hit(fg_engine *engine, const fg_emulsion *em, int64_t n_rays,
                                const float *origins, const float *dirs, int32_t *idx,
                                float *t0, float *t1);

/* Allocation-free transport of one tile.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops that forbid hipMalloc).  Same packets, draws and deposits as fg_expose on the
 * same emulsion, seed and configuration (the latent counts are identical), but every
 * array is the caller's device memory:
 *   nodes, records: the emulsion's threaded layout as fg_emulsion_get_device_layout
 *                   writes it (n_nodes * 32 and n_crystals * 48 bytes);
 *   counts:         n_crystals u32 latent counts in leaf order (record order), ADDED to;
 *   exposure:       tw * th f32 row-major map (NULL = uniform 1; required when
 *                   photons_per_pixel > 0).
 * em supplies the patch extents, optics and sizes (host data only; its device copy is
 * not used).  The only other device memory is a fixed FG_NOALLOC_SCRATCH_BYTES block per
 * GPU (work and statistics counters) that the library holds until the process ends; its
 * need does not grow with the tile.  Calls that overlap on one GPU share that block and
 * race.  stream: a hipStream_t, or NULL for the engine's stream.  stats NULL: the call
 * only enqueues (no host synchronisation, so it can be captured); stats given: it waits
 * for the kernel and reports its counters and time.  The call allocates nothing. */
#define FG_NOALLOC_SCRATCH_BYTES 256u
typedef struct fg_device_tile {
    const void *nodes;
    const void *records;
    uint32_t *counts;
    const float *exposure;
    uint32_t n_nodes, n_crystals;
    uint32_t tw, th;
} fg_device_tile;
fg_status fg_expose_noalloc(fg_engine *engine, const fg_emulsion *em, const fg_device_tile *tile,
                            const fg_exposure_config *cfg, uint64_t seed, void *stream,
                            fg_expose_stats *stats);

/* ------------------------------------------------------ drop-in tile render ----- */
/* Fill params the way image_pipeline.rs:455-515 and 924-976 do for a preset
 * ("tri-x", "tri-x-1600", "t-max", "hp5"), a tile of tw x th pixels at grain size
 * grain_size_px (pixel_size_um = 1 / grain_size_px), exposure and push. */
fg_status fg_tile_params_init(fg_tile_params *p, const char *preset, double grain_size_px,
                              uint32_t tw, uint32_t th, double exposure, double push,
                              uint64_t tile_seed);
/* GpuEngine::render_tile_all_gpu (gpu.rs:491-938): synthesis (seed), transport
 * (seed + 1), development, scan (seed + 2) of one tile; out_dneg receives
 * width_px * height_px D_neg values.  Zero pixels -> FG_OK with nothing written. */
fg_status fg_render_tile_all_gpu(fg_engine *engine, const float *tile_img,
                                 const fg_tile_params *params, float *out_dneg,
                                 fg_tile_stats *stats);

/* GpuEngine::render_tile_delta_d (gpu.rs:941-1228), the scan-only compatibility entry:
 * the n given crystals (pos, dims = semi-axes, quat = (w,i,j,k), morph 0 cubic / 1
 * tabular / 2 octahedral, fe = each crystal's filament expansion) are all developed; they
 * are scanned over the film_w x film_h x film_d patch at width_px x height_px with
 * samples_per_pixel rays per pixel, silver_opacity and seed params->seed, as
 * scan_emulsion does (transmittance.rs:65-171; D_neg = -log10 T + 0.05, no supersampling,
 * no OTF), and out_delta_d receives D_neg minus its 17-tap blur (the shader's literal
 * weights, which sum to 1.146; clamp-to-edge: gaussian.wgsl:14-59).  out_dneg (optional)
 * receives D_neg itself.
 * Zero pixels -> FG_OK with nothing written. */
fg_status fg_render_tile_delta_d(fg_engine *engine, int64_t n, const double *pos, const double *dims,
                                 const double *quat, const int32_t *morph, const double *fe,
                                 const fg_tile_params *params, float *out_delta_d, float *out_dneg);

/* ------------------------------------------------------------- frames ---------- */
#define FG_FRAME_PLAIN 0u   /* one launch set + synchronisation per tile */
#define FG_FRAME_BATCHED 1u /* persistent work-queue kernels over all tiles, replayed as
                               captured hipGraphs (transport; development; scan) */

/* A frame of width x height pixels cut into tile x tile tiles (edge tiles partial),
 * one emulsion per tile (seed per image_pipeline.rs:430-434), pixel units scaled by
 * pixel_size.  density_mult (per tile, NULL = 1) scales crystal density.
 * overlap > 0 lays the tiles out as the reference's tiled renderer does
 * (image_pipeline.rs:397-421): tile (tx, ty) spans [tx*tile - overlap, tx*tile + tile +
 * overlap) clipped to the frame, each with its own D_neg (fg_frame_read_tiles), blended
 * by fg_frame_read or fg_film_composite; it needs 2 * overlap <= tile and scan.otf = 0
 * (the scanner OTF then applies to the composited grain). */
typedef struct fg_frame_desc {
    uint32_t width, height, tile;
    uint32_t threads;             /* host threads for synthesis; 0 = all */
    double pixel_size;
    fg_emulsion_config emulsion;  /* width/height are set per tile */
    fg_exposure_config exposure;
    fg_development_config development;
    fg_scan_config scan;          /* width_px/height_px are set per tile */
    uint64_t seed;
    const float *exposure_map;    /* width * height, row-major, host */
    const float *density_mult;    /* tiles, or NULL */
    uint32_t synthesis;           /* 0 host Bridson (reference algorithm), 1 device dart
                                     throwing (fg_emulsion_synthesize_device) */
    uint32_t overlap;             /* tile overlap margin in pixels; 0 = abutting tiles */
} fg_frame_desc;

typedef struct fg_frame_stats {
    uint64_t tiles, crystals, developed, clumps, photons, rays, pixels;
    uint64_t bounces, nodes_visited, prims_tested, abs_events; /* transport counters */
    float ms_total, ms_transport, ms_develop, ms_scan;
    float ms_kernel_transport;    /* summed transport kernel time (HIP events) */
    float ms_kernel_scan;
    uint32_t graph_used;          /* 1: the batched render replayed its hipGraph; 0: plain
                                     mode, or direct launches after a failed capture */
    uint32_t pad_;
    uint64_t atomics;             /* global atomic additions of the deposits (batched) */
    uint64_t packets_run;         /* packets the batched transport started (= photons) */
} fg_frame_stats;

fg_status fg_frame_create(fg_engine *engine, const fg_frame_desc *desc, fg_frame **out);
void fg_frame_destroy(fg_frame *frame);
fg_status fg_frame_render(fg_engine *engine, fg_frame *frame, uint32_t mode,
                          fg_frame_stats *stats);
/* Frame D_neg (width * height, row-major) of the last render; with overlapping tiles
 * the Hann-weighted blend d_accum / max(weight, 1e-5) (image_pipeline.rs:593-654). */
fg_status fg_frame_read(fg_engine *engine, const fg_frame *frame, float *out);
/* D_neg of every tile of the last render, tiles in frame order, each tw * th row-major
 * (tw, th the tile's clipped extent); n = the total.  Returns the total in *n_total when
 * out is NULL. */
fg_status fg_frame_read_tiles(fg_engine *engine, const fg_frame *frame, int64_t n, float *out,
                              int64_t *n_total);
/* Per-tile latent counts of the last render, crystals in tile order then the tile's
 * crystal order; n must equal the frame's crystal count. */
fg_status fg_frame_read_counts(fg_engine *engine, const fg_frame *frame, int64_t n,
                               uint32_t *counts);
/* Development of the last render (developed flag, filament expansion, clump id or -1),
 * crystals in tile order then the tile's crystal order. */
/* Tile `tile` of a frame (row-major tile order): its crystal count n, its transport BVH's
 * node count n_nodes, its crystals in crystal order (pos, dims = semi-axes, quat = (w,i,j,k),
 * morph, sens; NULL to skip) and the device layout the transport reads (nodes: n_nodes
 * 32-byte threaded nodes; records: n 48-byte records in leaf order; NULL to skip), as
 * fg_emulsion_get_device_layout gives them for an emulsion.  For frames synthesised on
 * the device these come from the device build. */
fg_status fg_frame_get_tile(fg_engine *engine, const fg_frame *frame, int64_t tile, int64_t *n,
                            int64_t *n_nodes, double *pos, double *dims, double *quat, int32_t *morph,
                            double *sens, void *nodes, void *records);
fg_status fg_frame_read_development(fg_engine *engine, const fg_frame *frame, int64_t n,
                                    uint8_t *developed, double *fe, int64_t *clump_id);
int64_t fg_frame_crystal_count(const fg_frame *frame);

/* ---- differentiable (expected) model and its vector-Jacobian product ------------
 * No counterpart in the reference: the reference path is not differentiable (integer
 * latent counts, a Boolean development mask).  The expected model replaces each by its
 * expectation over the packet's stochastic rounding and the development threshold:
 *   dose_i  = sum over replayed packets and their hits on crystal i of
 *             5 * E(px) * w * (1 - exp(-mu_abs * L)) * sens_i
 *             (the mean of the latent count of transport.rs:117-134; packets replay
 *             the photon stream of fg_expose's stratified sampler without the rounding
 *             and roulette draws, and end once their relative weight falls below
 *             weight_cutoff);
 *   prob_i  = P[Poisson(dose_i) >= threshold]   (development.rs:69-84 in expectation);
 *   T(ray)  = prod over crystals hit, geometry scaled by the unboosted filament
 *             expansion, of 1 - prob_i * (1 - exp(-sigma * chord_i));
 *   out(px) = scan mode map of the mean T over the pixel's rays (transmittance.rs:
 *             137-152, image_pipeline.rs:1006-1011), rays as fg_scan's.
 * Gradients: exposure map and mu_abs, antihalation reflectance by path replay with
 * forward tangents of the packet weight; HG anisotropy g by the score function of the
 * Henyey-Greenstein density (the packet's path is held fixed). */
typedef struct fg_expected_model {
    fg_exposure_config exposure;       /* photons_per_pixel > 0; crystal_absorption_coeff */
    fg_optics optics;                  /* anisotropy_g, antihalation_reflectance */
    fg_development_config development; /* threshold_photons, filament_expansion */
    fg_scan_config scan;               /* otf must be 0 */
    uint64_t expose_seed, scan_seed;
    double weight_cutoff;              /* 0 = 1e-6 */
} fg_expected_model;

typedef struct fg_param_grads {
    double crystal_absorption_coeff;
    double anisotropy_g;
    double antihalation_reflectance;
    double reserved_;
} fg_param_grads;

typedef struct fg_adjoint_stats {
    uint64_t packets, bounces, events;
    float ms_replay;    /* path-replay kernel (dose or gradient pass) */
    float ms_scan;      /* expected scan or its adjoint */
    float ms_total;     /* the whole call, host included */
    float pad_;
    uint64_t atomics;   /* global f64 atomic additions of the dose pass after the wave
                           pre-reduction (events without it); 0 for the gradient pass */
} fg_adjoint_stats;

/* Expected-model forward on one tile: dose and prob per crystal (crystal order, each
 * optional) and out (scan.width_px * scan.height_px).  dose is what fg_vjp takes back. */
fg_status fg_expected_forward(fg_engine *engine, const fg_emulsion *em,
                              const fg_expected_model *model, const float *exposure_map,
                              uint32_t tw, uint32_t th, float *dose, float *prob, float *out,
                              fg_adjoint_stats *stats);
/* Vector-Jacobian product of fg_expected_forward: given grad_out = dL/d(out), writes
 * dL/d(exposure map) (tw * th), the parameter gradients and, optionally, dL/d(dose) per
 * crystal.  dose: the forward's output for the same inputs, or NULL to recompute it. */
fg_status fg_vjp(fg_engine *engine, const fg_emulsion *em, const fg_expected_model *model,
                 const float *exposure_map, uint32_t tw, uint32_t th, const float *dose,
                 const float *grad_out, float *grad_exposure, fg_param_grads *grads,
                 float *grad_dose, fg_adjoint_stats *stats);

/* ---- full-frame black-and-white film render ---------------------------------------
 * render_film_black_and_white_tiled (image_pipeline.rs:299-904) with every stage on the
 * GPU: panchromatic luma (:259-296), camera and gelatin optics (:133-252), overlapping
 * tiles rendered as one batched frame (:380-592, render_tile_all_gpu per tile), Hann
 * accumulation (:593-649), low-frequency separation (:651-706), sensitometric
 * compositing (:708-820), scanner OTF (:822-855) and the 8-bit output (:857-891).
 * Resampling to another output resolution and supersampling (:306-343) are the
 * image crate's filters and stay with the caller: the input is at output resolution. */
typedef struct fg_film_options {   /* FilmRenderOptions (image_pipeline.rs:20-92) */
    char preset[32];               /* "tri-x" (default), "tri-x-1600", "t-max", "hp5", ... */
    double grain_size_px;          /* 2.2 */
    double grain_strength;         /* 1.0 */
    double exposure;               /* 1.0 */
    double scale;                  /* 1.0 */
    double push;                   /* 0.0 */
    uint64_t seed;                 /* 42 */
    uint32_t tile_size;            /* 320 (>= 48) */
    uint32_t reference_res;        /* 6000 */
    uint32_t auto_scale;           /* 0 */
    uint32_t scanner_otf;          /* 1 */
    uint32_t synthesis;            /* 0 host Bridson, 1 device dart throwing */
    uint32_t frame_mode;           /* FG_FRAME_BATCHED (default) or FG_FRAME_PLAIN */
    uint32_t threads;              /* host threads for synthesis; 0 = all */
    uint32_t pad_;
} fg_film_options;

/* Sizes render_film_black_and_white_tiled derives (image_pipeline.rs:346-385, :651-658). */
typedef struct fg_film_plan {
    double grain_size_px;          /* after manual / automatic scaling, in [0.05, 15] */
    double strength;               /* effective grain strength */
    double push;                   /* effective push */
    double pixel_size_um;          /* 1 / grain_size_px */
    uint32_t tile_step, pad, tiles_x, tiles_y;
    float filter_sigma;            /* low-frequency Gaussian */
    uint32_t k_rad;
    float base_nominal;            /* per-preset grain amplitude before push */
    uint32_t pad_;
} fg_film_plan;

/* ImageProcessReport (image_pipeline.rs:95-105) with the real development counts of
 * every tile, the compositing statistics and per-stage times. */
typedef struct fg_film_report {
    uint32_t in_w, in_h, out_w, out_h;
    uint64_t total_tiles, total_crystals, developed_crystals, total_clumps;
    uint64_t photons, rays;
    double elapsed_seconds;
    float high_mean, high_std, norm_factor, pad_;
    float ms_luma_optics, ms_synthesis, ms_render, ms_composite;
    float ms_kernel_optics, ms_kernel_composite;   /* HIP events around the kernels */
    float ms_kernel_transport, ms_kernel_scan;
} fg_film_report;

/* Compositing statistics of fg_film_composite. */
typedef struct fg_composite_stats {
    float high_mean, high_std, norm_factor, filter_sigma;
    uint32_t k_rad, tiles_x, tiles_y, pad_;
    float ms_kernels;              /* HIP events around the compositing kernels */
    float pad2_;
} fg_composite_stats;

/* FilmRenderOptions::default() (image_pipeline.rs:72-92). */
fg_status fg_film_options_default(fg_film_options *opts);
/* Derived sizes for a width x height frame; host only, no device work. */
fg_status fg_film_plan_for(const fg_film_options *opts, uint32_t width, uint32_t height,
                           fg_film_plan *plan);
/* The whole render: pixels is width x height gray (channels 1) or interleaved RGB
 * (channels 3); out receives width x height 8-bit values. */
fg_status fg_render_film_bw(fg_engine *engine, const uint8_t *pixels, uint32_t width, uint32_t height,
                            uint32_t channels, const fg_film_options *opts, uint8_t *out,
                            fg_film_report *report);
/* Stage entry points (host buffers in and out): */
/* convert_to_panchromatic_luma of interleaved RGB for a preset's spectral weights;
 * ms_kernel (optional) receives the kernel's time (HIP events) */
fg_status fg_film_luma(fg_engine *engine, const uint8_t *rgb, uint32_t width, uint32_t height,
                       const char *preset, uint8_t *gray, float *ms_kernel);
/* apply_camera_and_gelatin_optics: h_effective (width * height f32) of a gray image;
 * ms_kernel (optional) receives the time of the gate and optics kernels (HIP events) */
fg_status fg_film_optics(fg_engine *engine, const uint8_t *gray, uint32_t width, uint32_t height,
                         float *h_effective, float *ms_kernel);
/* Compositing of rendered tiles for plan's layout: tiles = every tile's D_neg in frame
 * order (fg_frame_read_tiles' layout, n_tile_values in total), gray and h_effective at
 * width * height.  Optional outputs (NULL to skip), width * height f32 each: d_norm
 * (Hann blend), d_high (d_norm minus its Gaussian low-pass), grain (the grain field
 * after the scanner OTF when the plan's options enable it). */
fg_status fg_film_composite(fg_engine *engine, const fg_film_options *opts, uint32_t width,
                            uint32_t height, const uint8_t *gray, const float *h_effective,
                            const float *tiles, int64_t n_tile_values, uint8_t *out,
                            float *d_norm, float *d_high, float *grain, fg_composite_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* FGRAIN_HIP_H */
