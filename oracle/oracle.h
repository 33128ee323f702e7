/*
 * ORACLE -- test infrastructure only.
 *
 * A CPU restatement, in plain C, of the reference's photon-transport + development +
 * transmittance-scan path (GuyNguyen/fgrain, Rust).  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it, and only as the checker or the timed
 * CPU baseline; the product library never links or calls it.
 *
 * Every function names the reference file:line it restates.  FP64 functions follow the
 * Rust code (f64); the *_f32 twins evaluate the same formulas in float with the
 * operation order the GPU kernels use, for bit-exact comparisons on identical rays.
 *
 * Random streams: the reference draws from rand 0.8 StdRng (ChaCha12), which is not
 * vendored and cannot be built here, so stream-level parity with the reference is
 * unpinned.  The oracle instead uses Philox4x32-10 (Salmon et al., SC'11) with the
 * stream layouts below, which the GPU kernels share; distributions are pinned by
 * statistical and closed-form tests.
 */
#ifndef FGRAIN_ORACLE_H
#define FGRAIN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Crystal set, FP64 structure-of-arrays (HalideCrystal, crystal.rs:21-51). */
typedef struct {
    int64_t n;
    const double *pos;        /* n*3 centroid */
    const double *dims;       /* n*3 semi-axes */
    const double *quat;       /* n*4 unit quaternion (w, i, j, k) */
    const int32_t *morph;     /* n: 0 cubic, 1 tabular, 2 octahedral */
    const double *sens;       /* n sensitivity */
    const uint8_t *developed; /* n or NULL (= none developed) */
    const double *fe;         /* n filament expansion or NULL (= 1.0) */
} orc_crystals;

/* Flattened BVH exactly as bvh.rs:129-295 builds it (caller allocates 2n nodes). */
typedef struct {
    int64_t n_nodes;
    double *bmin, *bmax;      /* n_nodes*3 */
    uint32_t *left, *right, *offset, *count;
    uint32_t *indices;        /* n */
} orc_bvh_out;

typedef struct {
    double W, H, D;                 /* slab extents */
    double mu_a, mu_s, g, n_refr;   /* GelatinOptics (optics.rs:41-60) */
    double r_ahu, od_ahu;
    double mu_abs;                  /* ExposureConfig::crystal_absorption_coeff */
    double theta_max;               /* ExposureConfig::illumination_angle_rad */
    double photons_per_area;        /* sampler 0: photons per unit area at E=1 */
    int32_t max_bounces;
    int32_t sampler;                /* 0: reference chunked uniform; 1: photon-keyed (GPU) */
    int64_t n_photons;              /* sampler 1 with ppp == 0 */
    int32_t ppp;                    /* sampler 1: photons per pixel, stratified */
    int32_t threads;                /* worker threads; 0 = all online cpus */
    uint64_t seed;
    int32_t tw, th;                 /* exposure map size */
    int32_t pad_;
    double base_thickness;          /* extension: film base below the emulsion (0 = reference) */
} orc_expose_cfg;

typedef struct {
    int64_t launched, fresnel_rejected, zero_weight, bounces;
    int64_t nodes_tested, prims_tested, hits, abs_events;
    int64_t abs_total, chunks;
    int64_t bounces_live, nodes_live;  /* bounces / nodes of packets with non-zero weight */
} orc_expose_stats;

typedef struct {
    int64_t total, developed, fixed, clumps, max_clump;
    double avg_clump;
} orc_dev_report;

typedef struct {
    double W, H, D, r_min, mean_radius, size_sigma, aspect, base_sens, sens_sigma;
    double min_radius, min_half_thickness;
    int32_t morph, k_samples;
} orc_synth_cfg;

/* --- random streams ------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* --- geometry (crystal.rs, bvh.rs) ---------------------------------------------- */
int orc_crystal_intersect_f64(const orc_crystals *cs, int64_t i, const double o[3],
                              const double d[3], double t[2]);
int orc_crystal_intersect_f32(const orc_crystals *cs, int64_t i, const float o[3],
                              const float d[3], float t[2]);
int64_t orc_bvh_build(const orc_crystals *cs, orc_bvh_out *out);
/* All hits in the reference's DFS callback order for n_rays rays; hits of ray r are
 * written at [r*max_hits, ...) and their number to nhits[r].  f32 twin uses
 * outward-rounded float boxes and float crystal records. */
void orc_all_hits_f64(const orc_crystals *cs, int64_t n_rays, const double *o,
                      const double *d, int32_t max_hits, int32_t *nhits, int32_t *idx,
                      double *t0, double *t1);
void orc_all_hits_f32(const orc_crystals *cs, int64_t n_rays, const float *o,
                      const float *d, int32_t max_hits, int32_t *nhits, int32_t *idx,
                      float *t0, float *t1);
/* first hit per ray: smallest t_enter, ties to the earlier DFS hit; idx -1 = none */
void orc_first_hit_f32(const orc_crystals *cs, int64_t n_rays, const float *o, const float *d,
                       int32_t threads, int32_t *idx, float *t0, float *t1);
int64_t orc_query_radius(const orc_crystals *cs, const double c[3], double r,
                         int32_t *out, int64_t cap);

/* --- stages ------------------------------------------------------------------------ */
int orc_expose(const orc_crystals *cs, const float *exposure_map, const orc_expose_cfg *cfg,
               uint32_t *counts, orc_expose_stats *st);
int orc_develop(const orc_crystals *cs, const uint32_t *counts, uint32_t threshold,
                double fe, double pf, uint8_t *developed, double *fe_out, int64_t *clump_id,
                orc_dev_report *rep);
int orc_scan(const orc_crystals *cs, double W, double H, double D, int32_t w_px,
             int32_t h_px, int32_t spp, int32_t ss, double sigma, uint64_t seed,
             int32_t threads, float *T, int64_t *rays_nodes);
void orc_dneg(const float *T, int64_t n, float *D);
void orc_otf3(const float *in, int32_t w, int32_t h, float *out);
double orc_free_path(double mu_t, double xi);
void orc_hg(const double d[3], double g, double xi1, double xi2, double out[3]);
double orc_fresnel(double n, double cos_i);
int orc_antihalation(double r_ahu, double od, uint64_t seed, double *energy, double d[3]);
int64_t orc_synthesize(const orc_synth_cfg *cfg, uint64_t seed, int64_t cap, double *pos,
                       double *dims, double *quat, int32_t *morph, double *sens);
/* twin of the library's device synthesis (parallel dart throwing, per-crystal streams);
 * no reference counterpart */
int64_t orc_darts(const orc_synth_cfg *cfg, uint64_t seed, int64_t cap, double *pos, double *dims,
                  double *quat, int32_t *morph, double *sens);

/* --- expected (differentiable) model: no reference counterpart ---------------------- */
typedef struct {
    orc_expose_cfg x;       /* sampler 1 with ppp > 0 and an exposure map; r_ahu = R */
    double cut;             /* relative weight cutoff of a replayed packet (0 = 1e-6) */
    double fe;              /* filament expansion of the expected-scan geometry */
    double sigma;           /* silver opacity */
    double pf;              /* clumping proximity factor of the touch graph */
    int32_t w_px, h_px, spp, ss;
    uint64_t scan_seed;
} orc_expected_cfg;

/* expected dose per crystal (crystal order); bounces_events[2] optional */
int orc_expected_dose(const orc_crystals *cs, const float *map, const orc_expected_cfg *e,
                      double *dose, int64_t *bounces_events);
/* P[Poisson(dose) >= thr] and d/d(dose) */
void orc_poisson_tail(const double *dose, int64_t n, uint32_t thr, double *p, double *dp);
/* semi-axes times fe * boost per crystal; gT == NULL: T (w_px * h_px mean ray products);
 * else gprob, gboost (n) for dL/dT = gT */
int orc_expected_scan(const orc_crystals *cs, const double *prob, const double *boost,
                      const orc_expected_cfg *e, const double *gT, double *T, double *gprob,
                      double *gboost);
/* touch graph at unboosted expansion (CSR); returns entries, or -entries if cap is short */
int64_t orc_touch_graph(const orc_crystals *cs, double fe, double pf, int64_t *ptr, int32_t *idx,
                        int64_t cap);
/* chord of crystal i scaled by s about its centre and d chord / d s */
int orc_chord_scaled(const orc_crystals *cs, int64_t i, double s, const double o[3],
                     const double d[3], double out[2]);
/* gradient replay: gmap (tw * th) = dL/dE, grads = dL/d(mu_abs, g, R) for dL/ddose = gdose */
int orc_replay_grad(const orc_crystals *cs, const float *map, const orc_expected_cfg *e,
                    const double *gdose, double *gmap, double grads[3]);
double orc_hg_score(double g, double mu);

/* --- full-frame black-and-white compositing (film.c, image_pipeline.rs:133-891) ------ */
typedef struct {
    double grain_size_px;   /* after manual and automatic scaling */
    double strength;        /* effective grain strength */
    uint32_t tile_step, pad, tiles_x, tiles_y;
} orc_film_plan_out;

typedef struct {
    int32_t width, height, tile_step, pad;
    double grain_size_px;
    float strength;         /* effective grain strength */
    float base_nominal;     /* per preset: t-max 0.024, hp5 0.032, tri-x-1600 0.042, else 0.038 */
    float push;             /* effective push */
    int32_t scanner_otf;
    int32_t fold;           /* 0 sequential f32 statistics (reference), 1 f64 sums */
} orc_film_cfg;

typedef struct {
    float high_mean, high_std, norm_factor, filter_sigma;
    int32_t k_rad, tiles_x, tiles_y, pad_;
} orc_film_stats;

void orc_film_luma(const uint8_t *rgb, int64_t n, float rw, float gw, float bw, uint8_t *out);
int orc_film_plan(double grain_size_px, double grain_strength, double scale, int32_t auto_scale,
                  uint32_t reference_res, uint32_t tile_size, uint32_t out_w, uint32_t out_h,
                  orc_film_plan_out *p);
void orc_film_optics(const uint8_t *src, int32_t w, int32_t h, float *h_eff, int32_t margins[4]);
int64_t orc_film_tile_values(int32_t w, int32_t h, int32_t step, int32_t pad);
int orc_film_composite(const orc_film_cfg *c, const uint8_t *src, const float *h_eff, const float *tiles,
                       uint8_t *out, float *d_norm, float *d_low, float *grain, float *scanned,
                       float *tone, orc_film_stats *st);

#ifdef __cplusplus
}
#endif
#endif
