/*
 * ORACLE -- test infrastructure only (see oracle.h).  CPU restatement of the
 * reference's hot path; never linked into the product library.
 *
 * Reference files restated (GuyNguyen/fgrain):
 *   crystal.rs:106-230   AABB rule and ray/crystal intersection
 *   bvh.rs:21-402        AABB, median-split build, DFS all-hits traversal, radius query
 *   optics.rs:79-196     free path, Henyey-Greenstein, Schlick Fresnel, anti-halation
 *   transport.rs:48-188  Monte-Carlo exposure (Rayon chunking, Beer-Lambert, RR)
 *   development.rs:59-188 threshold + BFS clumping + clump boost
 *   transmittance.rs:65-171, image_pipeline.rs:990-1014  scan + D_neg
 *   image_pipeline.rs:822-855  scanner OTF 3-tap
 *   emulsion.rs:155-220, poisson.rs:14-132  emulsion synthesis (Bridson 3D)
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define TAU 6.283185307179586
#define PI_ 3.141592653589793

/* Philox counter layouts shared with the GPU kernels:
 *   photon stream:  ctr = (lo32(photon), hi32(photon), draw >> 2, TAG_PHOTON), word draw & 3
 *   scan jitter:    ctr = (pixel, ray >> 1, TAG_SCAN, 0), words 2*(ray&1) (x), +1 (y)
 *   sequential:     ctr = (lo32(block), hi32(block), 0, TAG_SEQ), u64 = lo | hi << 32 */
#define TAG_PHOTON 0x50484f54u
#define TAG_SCAN 0x5343414eu
#define TAG_SEQ 0x53455120u
#define TAG_DART 0x44415254u
#define TAG_ATTR 0x41545452u

/* ---------------------------------------------------------------- Philox4x32-10 -- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

typedef struct {
    int mode;         /* 0 sequential (f64 uniforms), 1 photon-keyed (f32-word uniforms),
                         2 per-crystal stream of device synthesis (f64 uniforms) */
    uint32_t key[2];
    uint64_t ctr;     /* mode 0: next block; mode 1: photon id */
    uint32_t k;       /* mode 1: draw index */
    uint32_t blk;     /* mode 1: cached block index */
    int valid;
    int pos;          /* mode 0: next word of buf */
    uint32_t buf[4];
} orng;

static void rng_seq(orng *r, uint64_t seed)
{
    memset(r, 0, sizeof(*r));
    r->mode = 0;
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->pos = 4;
}

static void rng_photon(orng *r, uint64_t seed, uint64_t photon)
{
    memset(r, 0, sizeof(*r));
    r->mode = 1;
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->ctr = photon;
}

static void rng_point(orng *r, uint64_t seed, uint64_t point)
{
    memset(r, 0, sizeof(*r));
    r->mode = 2;
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->ctr = point;
    r->pos = 4;
}

static uint32_t seq_u32(orng *r)
{
    if (r->pos == 4) {
        if (r->mode == 2) {
            /* ctr = (lo32(point), hi32(point), block, TAG_ATTR) */
            const uint32_t c[4] = {(uint32_t)r->ctr, (uint32_t)(r->ctr >> 32), r->k++, TAG_ATTR};
            orc_philox4x32_10(c, r->key, r->buf);
        } else {
            const uint32_t c[4] = {(uint32_t)r->ctr, (uint32_t)(r->ctr >> 32), 0u, TAG_SEQ};
            orc_philox4x32_10(c, r->key, r->buf);
            r->ctr++;
        }
        r->pos = 0;
    }
    return r->buf[r->pos++];
}

static uint64_t seq_u64(orng *r)
{
    const uint64_t lo = seq_u32(r);
    const uint64_t hi = seq_u32(r);
    return lo | (hi << 32);
}

static uint32_t photon_word(orng *r)
{
    const uint32_t b = r->k >> 2;
    if (!r->valid || b != r->blk) {
        const uint32_t c[4] = {(uint32_t)r->ctr, (uint32_t)(r->ctr >> 32), b, TAG_PHOTON};
        orc_philox4x32_10(c, r->key, r->buf);
        r->blk = b;
        r->valid = 1;
    }
    return r->buf[r->k++ & 3u];
}

/* uniform in [0,1): 53-bit for sequential streams, 24-bit (exact float) for photon streams */
static double rng_u(orng *r)
{
    if (r->mode != 1) return (double)(seq_u64(r) >> 11) * 0x1p-53;
    return (double)((float)(photon_word(r) >> 8) * 0x1p-24f);
}

/* integer in [0, n) by 64x64->128 multiply-high (the spec for gen_range(0..n)) */
static uint64_t rng_below(orng *r, uint64_t n)
{
    return (uint64_t)(((unsigned __int128)seq_u64(r) * n) >> 64);
}

/* ------------------------------------------------- precision-generic geometry -- */
#define R double
#define FN(x) x##_d
#define RMIN fmin
#define RMAX fmax
#define RSQRT sqrt
#include "orc_geom_impl.h"
#undef R
#undef FN
#undef RMIN
#undef RMAX
#undef RSQRT
#define R float
#define FN(x) x##_s
#define RMIN fminf
#define RMAX fmaxf
#define RSQRT sqrtf
#include "orc_geom_impl.h"
#undef R
#undef FN
#undef RMIN
#undef RMAX
#undef RSQRT

/* crystal record with expanded semi-axes (crystal.rs:145-156) */
typedef struct {
    double c[3], q[4], dims[3], sens;
    int morph;
} ocrys;
typedef struct {
    float c[3], q[4], dims[3], sens;
    int morph;
} ocrys32;

static double expansion_of(const orc_crystals *cs, int64_t i)
{
    if (cs->developed && cs->developed[i]) return cs->fe ? cs->fe[i] : 1.0;
    return 1.0;
}

static void make_rec(const orc_crystals *cs, int64_t i, ocrys *o)
{
    const double e = expansion_of(cs, i);
    for (int a = 0; a < 3; a++) {
        o->c[a] = cs->pos[3 * i + a];
        o->dims[a] = cs->dims[3 * i + a] * e;
    }
    for (int a = 0; a < 4; a++) o->q[a] = cs->quat[4 * i + a];
    o->sens = cs->sens ? cs->sens[i] : 1.0;
    o->morph = cs->morph[i];
}

static void make_rec32(const orc_crystals *cs, int64_t i, ocrys32 *o)
{
    ocrys r;
    make_rec(cs, i, &r);
    for (int a = 0; a < 3; a++) { o->c[a] = (float)r.c[a]; o->dims[a] = (float)r.dims[a]; }
    for (int a = 0; a < 4; a++) o->q[a] = (float)r.q[a];
    o->sens = (float)r.sens;
    o->morph = r.morph;
}

int orc_crystal_intersect_f64(const orc_crystals *cs, int64_t i, const double o[3],
                              const double d[3], double t[2])
{
    ocrys r;
    make_rec(cs, i, &r);
    return crystal_intersect_d(r.c, r.q, r.dims, r.morph, o, d, &t[0], &t[1]);
}

int orc_crystal_intersect_f32(const orc_crystals *cs, int64_t i, const float o[3],
                              const float d[3], float t[2])
{
    ocrys32 r;
    make_rec32(cs, i, &r);
    return crystal_intersect_s(r.c, r.q, r.dims, r.morph, o, d, &t[0], &t[1]);
}

/* ------------------------------------------------------------------------ BVH -- */
typedef struct {
    double mn[3], mx[3];
    uint32_t left, right, offset, count;
} onode;

typedef struct {
    onode *nodes;
    int64_t nn;
    uint32_t *idx;
    int64_t n;
} obvh;

/* HalideCrystal::aabb (crystal.rs:106-133): centre +- max semi-axis * expansion, for
 * every orientation (not conservative for rotated cubes). */
static void crystal_aabb(const orc_crystals *cs, int64_t i, int use_dev, double mn[3],
                         double mx[3])
{
    const double e = use_dev ? expansion_of(cs, i) : 1.0;
    const double *dm = cs->dims + 3 * i;
    const double m = fmax(fmax(dm[0], dm[1]), dm[2]) * e;
    for (int a = 0; a < 3; a++) {
        mn[a] = cs->pos[3 * i + a] - m;
        mx[a] = cs->pos[3 * i + a] + m;
    }
}

typedef struct {
    obvh *b;
    const double *amn, *amx, *cen;
} build_ctx;

/* BVH::subdivide (bvh.rs:166-295) */
static uint32_t subdivide(build_ctx *X, int64_t start, int64_t count)
{
    obvh *b = X->b;
    const uint32_t node_idx = (uint32_t)b->nn;
    double bmn[3] = {INFINITY, INFINITY, INFINITY}, bmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    double cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t k = start; k < start + count; k++) {
        const uint32_t id = b->idx[k];
        for (int a = 0; a < 3; a++) {
            bmn[a] = fmin(bmn[a], X->amn[3 * id + a]);
            bmx[a] = fmax(bmx[a], X->amx[3 * id + a]);
            cmn[a] = fmin(cmn[a], X->cen[3 * id + a]);
            cmx[a] = fmax(cmx[a], X->cen[3 * id + a]);
        }
    }
    onode *nd = &b->nodes[b->nn++];
    memcpy(nd->mn, bmn, sizeof bmn);
    memcpy(nd->mx, bmx, sizeof bmx);
    nd->left = nd->right = 0;
    if (count <= 4) {
        nd->offset = (uint32_t)start;
        nd->count = (uint32_t)count;
        return node_idx;
    }
    const double ex = cmx[0] - cmn[0], ey = cmx[1] - cmn[1], ez = cmx[2] - cmn[2];
    const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
    const double ext[3] = {ex, ey, ez};
    if (ext[axis] <= 1e-7) {
        nd->offset = (uint32_t)start;
        nd->count = (uint32_t)count;
        return node_idx;
    }
    const double mid = 0.5 * (cmn[axis] + cmx[axis]);
    int64_t i = start, j = start + count - 1;
    while (i <= j) {
        const uint32_t id = b->idx[i];
        if (X->cen[3 * id + axis] < mid) {
            i++;
        } else {
            const uint32_t t = b->idx[i];
            b->idx[i] = b->idx[j];
            b->idx[j] = t;
            if (j == 0) break;
            j--;
        }
    }
    const int64_t left_count = i - start;
    const int64_t split = (left_count == 0 || left_count == count) ? count / 2 : left_count;
    nd->offset = 0;
    nd->count = 0;
    const uint32_t l = subdivide(X, start, split);
    const uint32_t r = subdivide(X, start + split, count - split);
    b->nodes[node_idx].left = l;
    b->nodes[node_idx].right = r;
    return node_idx;
}

/* BVH::build (bvh.rs:157-195) over the crystals selected by sel (NULL = all). */
static void bvh_build(const orc_crystals *cs, int use_dev, obvh *b)
{
    const int64_t n = cs->n;
    b->n = n;
    b->nn = 0;
    b->nodes = NULL;
    b->idx = NULL;
    if (n == 0) return;
    b->nodes = (onode *)malloc(sizeof(onode) * (size_t)(2 * n));
    b->idx = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    double *amn = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *amx = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *cen = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        b->idx[i] = (uint32_t)i;
        crystal_aabb(cs, i, use_dev, amn + 3 * i, amx + 3 * i);
        for (int a = 0; a < 3; a++) cen[3 * i + a] = 0.5 * (amn[3 * i + a] + amx[3 * i + a]);
    }
    build_ctx X = {b, amn, amx, cen};
    subdivide(&X, 0, n);
    free(amn);
    free(amx);
    free(cen);
}

static void bvh_free(obvh *b)
{
    free(b->nodes);
    free(b->idx);
}

int64_t orc_bvh_build(const orc_crystals *cs, orc_bvh_out *out)
{
    obvh b;
    bvh_build(cs, 1, &b);
    for (int64_t k = 0; k < b.nn; k++) {
        for (int a = 0; a < 3; a++) {
            out->bmin[3 * k + a] = b.nodes[k].mn[a];
            out->bmax[3 * k + a] = b.nodes[k].mx[a];
        }
        out->left[k] = b.nodes[k].left;
        out->right[k] = b.nodes[k].right;
        out->offset[k] = b.nodes[k].offset;
        out->count[k] = b.nodes[k].count;
    }
    for (int64_t i = 0; i < b.n; i++) out->indices[i] = b.idx[i];
    out->n_nodes = b.nn;
    bvh_free(&b);
    return out->n_nodes;
}

/* float box rounded outward from the double box, the layout the GPU traverses */
static float round_down(double x)
{
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}
static float round_up(double x)
{
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}

/* --------------------------------------------------------- DFS traversal -------- */
/* BVH::intersect_ray_all (bvh.rs:300-347): infinite ray, 64-entry stack, left pushed at
 * sp and right at sp+1 so the right child pops first; children are dropped when
 * sp + 2 >= 64.  Callback order = hit order. */
typedef struct {
    int64_t nodes, prims;
} tcount;

#define DFS_TRAVERSE(B, HITFN, PRIMFN, CTX, CNT)                                      \
    do {                                                                              \
        uint32_t stack_[64];                                                          \
        int sp_ = 0;                                                                  \
        stack_[sp_++] = 0;                                                            \
        while (sp_ > 0) {                                                             \
            const onode *nd_ = &(B)->nodes[stack_[--sp_]];                            \
            (CNT)->nodes++;                                                           \
            if (HITFN(nd_, CTX)) {                                                    \
                if (nd_->count > 0) {                                                 \
                    for (uint32_t k_ = nd_->offset; k_ < nd_->offset + nd_->count; k_++) { \
                        (CNT)->prims++;                                               \
                        PRIMFN((B)->idx[k_], CTX);                                    \
                    }                                                                 \
                } else if (sp_ + 2 < 64) {                                            \
                    stack_[sp_] = nd_->left;                                          \
                    stack_[sp_ + 1] = nd_->right;                                     \
                    sp_ += 2;                                                         \
                }                                                                     \
            }                                                                         \
        }                                                                             \
    } while (0)

typedef struct {
    const ocrys *cr;
    double o[3], d[3], inv[3];
    int32_t n, cap;
    int32_t *idx;
    double *t0, *t1;
} allhits_ctx;

static inline int ah_box(const onode *nd, allhits_ctx *c)
{
    return box_hit_d(nd->mn, nd->mx, c->o, c->inv);
}
static inline void ah_prim(uint32_t id, allhits_ctx *c)
{
    const ocrys *r = &c->cr[id];
    double a, b;
    if (crystal_intersect_d(r->c, r->q, r->dims, r->morph, c->o, c->d, &a, &b)) {
        if (c->n < c->cap) {
            c->idx[c->n] = (int32_t)id;
            c->t0[c->n] = a;
            c->t1[c->n] = b;
        }
        c->n++;
    }
}

static ocrys *make_recs(const orc_crystals *cs)
{
    ocrys *cr = (ocrys *)malloc(sizeof(ocrys) * (size_t)(cs->n ? cs->n : 1));
    for (int64_t i = 0; i < cs->n; i++) make_rec(cs, i, &cr[i]);
    return cr;
}

void orc_all_hits_f64(const orc_crystals *cs, int64_t n_rays, const double *o,
                      const double *d, int32_t max_hits, int32_t *nhits, int32_t *idx,
                      double *t0, double *t1)
{
    obvh b;
    bvh_build(cs, 1, &b);
    ocrys *cr = make_recs(cs);
    for (int64_t r = 0; r < n_rays; r++) {
        allhits_ctx c;
        c.cr = cr;
        for (int a = 0; a < 3; a++) {
            c.o[a] = o[3 * r + a];
            c.d[a] = d[3 * r + a];
            c.inv[a] = 1.0 / c.d[a];
        }
        c.n = 0;
        c.cap = max_hits;
        c.idx = idx + r * max_hits;
        c.t0 = t0 + r * max_hits;
        c.t1 = t1 + r * max_hits;
        tcount cnt = {0, 0};
        if (b.nn > 0) DFS_TRAVERSE(&b, ah_box, ah_prim, &c, &cnt);
        nhits[r] = c.n;
    }
    free(cr);
    bvh_free(&b);
}

/* float twin */
typedef struct {
    float mn[3], mx[3];
} fbox;

typedef struct {
    const ocrys32 *cr;
    const fbox *fb;
    const onode *base;
    float o[3], d[3], inv[3];
    int32_t n, cap;
    int32_t *idx;
    float *t0, *t1;
} allhits32_ctx;

static inline int ah32_box(const onode *nd, allhits32_ctx *c)
{
    const fbox *f = &c->fb[nd - c->base];
    return box_hit_s(f->mn, f->mx, c->o, c->inv);
}
static inline void ah32_prim(uint32_t id, allhits32_ctx *c)
{
    const ocrys32 *r = &c->cr[id];
    float a, b;
    if (crystal_intersect_s(r->c, r->q, r->dims, r->morph, c->o, c->d, &a, &b)) {
        if (c->n < c->cap) {
            c->idx[c->n] = (int32_t)id;
            c->t0[c->n] = a;
            c->t1[c->n] = b;
        }
        c->n++;
    }
}

void orc_all_hits_f32(const orc_crystals *cs, int64_t n_rays, const float *o,
                      const float *d, int32_t max_hits, int32_t *nhits, int32_t *idx,
                      float *t0, float *t1)
{
    obvh b;
    bvh_build(cs, 1, &b);
    ocrys32 *cr = (ocrys32 *)malloc(sizeof(ocrys32) * (size_t)(cs->n ? cs->n : 1));
    for (int64_t i = 0; i < cs->n; i++) make_rec32(cs, i, &cr[i]);
    fbox *fb = (fbox *)malloc(sizeof(fbox) * (size_t)(b.nn ? b.nn : 1));
    for (int64_t k = 0; k < b.nn; k++)
        for (int a = 0; a < 3; a++) {
            fb[k].mn[a] = round_down(b.nodes[k].mn[a]);
            fb[k].mx[a] = round_up(b.nodes[k].mx[a]);
        }
    for (int64_t r = 0; r < n_rays; r++) {
        allhits32_ctx c;
        c.cr = cr;
        c.fb = fb;
        c.base = b.nodes;
        for (int a = 0; a < 3; a++) {
            c.o[a] = o[3 * r + a];
            c.d[a] = d[3 * r + a];
            c.inv[a] = 1.0f / c.d[a];
        }
        c.n = 0;
        c.cap = max_hits;
        c.idx = idx + r * max_hits;
        c.t0 = t0 + r * max_hits;
        c.t1 = t1 + r * max_hits;
        tcount cnt = {0, 0};
        if (b.nn > 0) DFS_TRAVERSE(&b, ah32_box, ah32_prim, &c, &cnt);
        nhits[r] = c.n;
    }
    free(fb);
    free(cr);
    bvh_free(&b);
}

/* First hit of each ray (the GPU's fg_traverse_first_hit): the smallest t_enter over
 * intersect_ray_all's hits (bvh.rs:300-347, crystal.rs:140-217), ties to the earlier hit
 * in depth-first callback order; f32 twin of the device arithmetic.  Threads split the
 * rays (threads <= 0: one). */
typedef struct {
    allhits32_ctx base;
    int32_t best;
    float b0, b1;
} first32_ctx;

static inline int fh32_box(const onode *nd, first32_ctx *c) { return ah32_box(nd, &c->base); }
static inline void fh32_prim(uint32_t id, first32_ctx *c)
{
    const ocrys32 *r = &c->base.cr[id];
    float a, b;
    if (crystal_intersect_s(r->c, r->q, r->dims, r->morph, c->base.o, c->base.d, &a, &b) &&
        (c->best < 0 || a < c->b0)) {
        c->best = (int32_t)id;
        c->b0 = a;
        c->b1 = b;
    }
}

typedef struct {
    const obvh *b;
    const ocrys32 *cr;
    const fbox *fb;
    const float *o, *d;
    int64_t r0, r1;
    int32_t *idx;
    float *t0, *t1;
} first32_job;

static void *first32_worker(void *arg)
{
    first32_job *j = (first32_job *)arg;
    for (int64_t r = j->r0; r < j->r1; r++) {
        first32_ctx c;
        c.base.cr = j->cr;
        c.base.fb = j->fb;
        c.base.base = j->b->nodes;
        for (int a = 0; a < 3; a++) {
            c.base.o[a] = j->o[3 * r + a];
            c.base.d[a] = j->d[3 * r + a];
            c.base.inv[a] = 1.0f / c.base.d[a];
        }
        c.best = -1;
        c.b0 = c.b1 = 0.0f;
        tcount cnt = {0, 0};
        if (j->b->nn > 0) DFS_TRAVERSE(j->b, fh32_box, fh32_prim, &c, &cnt);
        j->idx[r] = c.best;
        j->t0[r] = c.b0;
        j->t1[r] = c.b1;
    }
    return NULL;
}

void orc_first_hit_f32(const orc_crystals *cs, int64_t n_rays, const float *o, const float *d,
                       int32_t threads, int32_t *idx, float *t0, float *t1)
{
    obvh b;
    bvh_build(cs, 1, &b);
    ocrys32 *cr = (ocrys32 *)malloc(sizeof(ocrys32) * (size_t)(cs->n ? cs->n : 1));
    for (int64_t i = 0; i < cs->n; i++) make_rec32(cs, i, &cr[i]);
    fbox *fb = (fbox *)malloc(sizeof(fbox) * (size_t)(b.nn ? b.nn : 1));
    for (int64_t k = 0; k < b.nn; k++)
        for (int a = 0; a < 3; a++) {
            fb[k].mn[a] = round_down(b.nodes[k].mn[a]);
            fb[k].mx[a] = round_up(b.nodes[k].mx[a]);
        }
    const int nt = threads > 1 ? threads : 1;
    first32_job *jobs = (first32_job *)malloc(sizeof(first32_job) * (size_t)nt);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
    for (int t = 0; t < nt; t++) {
        jobs[t] = (first32_job){&b, cr, fb, o, d, n_rays * t / nt, n_rays * (t + 1) / nt, idx, t0, t1};
        pthread_create(&th[t], NULL, first32_worker, &jobs[t]);
    }
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    free(fb);
    free(cr);
    bvh_free(&b);
}

/* BVH::query_radius (bvh.rs:350-402): inclusive box overlap, centre distance <= r. */
static int64_t query_radius(const obvh *b, const orc_crystals *cs, const double c[3],
                            double r, int32_t *out, int64_t cap)
{
    if (b->nn == 0) return 0;
    const double r2 = r * r;
    const double qmn[3] = {c[0] - r, c[1] - r, c[2] - r};
    const double qmx[3] = {c[0] + r, c[1] + r, c[2] + r};
    uint32_t stack[64];
    int sp = 0;
    int64_t n = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const onode *nd = &b->nodes[stack[--sp]];
        const int ov = nd->mn[0] <= qmx[0] && nd->mx[0] >= qmn[0] && nd->mn[1] <= qmx[1] &&
                       nd->mx[1] >= qmn[1] && nd->mn[2] <= qmx[2] && nd->mx[2] >= qmn[2];
        if (!ov) continue;
        if (nd->count > 0) {
            for (uint32_t k = nd->offset; k < nd->offset + nd->count; k++) {
                const uint32_t id = b->idx[k];
                const double dx = cs->pos[3 * id] - c[0], dy = cs->pos[3 * id + 1] - c[1],
                             dz = cs->pos[3 * id + 2] - c[2];
                if (dx * dx + dy * dy + dz * dz <= r2) {
                    if (n < cap) out[n] = (int32_t)id;
                    n++;
                }
            }
        } else if (sp + 2 < 64) {
            stack[sp] = nd->left;
            stack[sp + 1] = nd->right;
            sp += 2;
        }
    }
    return n;
}

int64_t orc_query_radius(const orc_crystals *cs, const double c[3], double r, int32_t *out,
                         int64_t cap)
{
    obvh b;
    bvh_build(cs, 1, &b);
    const int64_t n = query_radius(&b, cs, c, r, out, cap);
    bvh_free(&b);
    return n;
}

/* ------------------------------------------------------------------ transport -- */
typedef struct {
    const obvh *b;
    const ocrys *cr;
    const orc_expose_cfg *cfg;
    const float *map;
    uint32_t *counts;
    double mu_t, albedo, r0, halation_eff;
    int64_t n_total;    /* sampler 1 photon count */
    int64_t chunks, per_chunk;
    int64_t next;       /* work counter (atomic) */
    int64_t grain;
} texp;

typedef struct {
    orc_expose_stats st;
} tstat;

/* nearest-pixel exposure lookup of the tile path (image_pipeline.rs:949-961) */
static double exposure_at(const texp *X, double x, double y)
{
    const orc_expose_cfg *c = X->cfg;
    if (!X->map) return 1.0;
    const double u = fmin(fmax(x / c->W, 0.0), 1.0);
    const double v = fmin(fmax(y / c->H, 0.0), 1.0);
    double fx = round(u * (double)(c->tw - 1)), fy = round(v * (double)(c->th - 1));
    int64_t px = (int64_t)fx, py = (int64_t)fy;
    if (px > c->tw - 1) px = c->tw - 1;
    if (py > c->th - 1) py = c->th - 1;
    if (px < 0) px = 0;
    if (py < 0) py = 0;
    return (double)X->map[py * c->tw + px];
}

static void normalize3(double v[3])
{
    const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= n;
    v[1] /= n;
    v[2] /= n;
}

typedef struct {
    const texp *X;
    orng *rng;
    const double *o, *d;
    double inv[3];
    double step, w;
    orc_expose_stats *st;
} hitctx;

static inline int tr_box(const onode *nd, hitctx *h)
{
    return box_hit_d(nd->mn, nd->mx, h->o, h->inv);
}

/* transport.rs:112-137: Beer-Lambert on every hit starting before the step end, in
 * traversal order; latent photons by floor + Bernoulli(frac); weight attenuation. */
static inline void tr_prim(uint32_t id, hitctx *h)
{
    const ocrys *r = &h->X->cr[id];
    double t0, t1;
    if (!crystal_intersect_d(r->c, r->q, r->dims, r->morph, h->o, h->d, &t0, &t1)) return;
    h->st->hits++;
    if (!(t0 < h->step)) return;
    const double L = fmax(fmin(t1, h->step) - t0, 0.0);
    if (!(L > 0.0)) return;
    const double pa = 1.0 - exp(-h->X->cfg->mu_abs * L);
    const double absorbed = h->w * pa * r->sens;
    const double expected = absorbed * 5.0;
    double fl = floor(expected);
    uint32_t whole;
    if (!(fl > 0.0)) whole = 0;                 /* saturating `as u32`, NaN -> 0 */
    else if (fl >= 4294967295.0) whole = 4294967295u;
    else whole = (uint32_t)fl;
    const double frac = expected - (double)whole;
    const double p = fmin(fmax(frac, 0.0), 1.0);
    const uint32_t extra = (rng_u(h->rng) < p) ? 1u : 0u;
    const uint32_t count = whole + extra;
    if (count > 0) {
        __atomic_fetch_add(&h->X->counts[id], count, __ATOMIC_RELAXED);
        h->st->abs_events++;
        h->st->abs_total += count;
    }
    h->w *= fmax(1.0 - pa * 0.9, 0.05);
}

/* optics.rs:113-143; returns the sampled cosine */
static double sample_hg(double d[3], double g, orng *rng)
{
    const double xi1 = rng_u(rng), xi2 = rng_u(rng);
    double ct;
    if (fabs(g) < 1e-4) {
        ct = 2.0 * xi1 - 1.0;
    } else {
        const double term = (1.0 - g * g) / (1.0 - g + 2.0 * g * xi1);
        ct = (1.0 + g * g - term * term) / (2.0 * g);
    }
    ct = fmin(fmax(ct, -1.0), 1.0);
    const double stt = sqrt(fmax(1.0 - ct * ct, 0.0));
    const double phi = TAU * xi2;
    double w[3] = {d[0], d[1], d[2]};
    normalize3(w);
    double u[3];
    if (fabs(w[0]) > 0.1) { /* y x w */
        u[0] = 1.0 * w[2] - 0.0 * w[1];
        u[1] = 0.0 * w[0] - 0.0 * w[2];
        u[2] = 0.0 * w[1] - 1.0 * w[0];
    } else {                /* x x w */
        u[0] = 0.0 * w[2] - 0.0 * w[1];
        u[1] = 0.0 * w[0] - 1.0 * w[2];
        u[2] = 1.0 * w[1] - 0.0 * w[0];
    }
    normalize3(u);
    const double v[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2],
                         w[0] * u[1] - w[1] * u[0]};
    const double a = stt * cos(phi), bb = stt * sin(phi);
    for (int k = 0; k < 3; k++) d[k] = (u[k] * a + v[k] * bb) + w[k] * ct;
    normalize3(d);
    return ct;
}

/* optics.rs:160-196; returns 0 when the packet dies */
static int antihalation(const texp *X, double *w, double d[3], orng *rng)
{
    *w *= X->halation_eff;
    if (*w < 1e-4) {
        if (rng_u(rng) < 0.1) *w /= 0.1;
        else return 0;
    }
    if (rng_u(rng) < 0.4) {
        const double u1 = rng_u(rng), u2 = rng_u(rng);
        const double r = sqrt(u1), phi = TAU * u2;
        d[0] = r * cos(phi);
        d[1] = r * sin(phi);
        d[2] = -sqrt(fmax(1.0 - u1, 0.0));
        normalize3(d);
    } else {
        d[2] = -fabs(d[2]);
        normalize3(d);
    }
    return 1;
}

static double fresnel(const texp *X, double cos_i)
{
    const double c = fmin(fmax(fabs(cos_i), 0.0), 1.0);
    const double m = 1.0 - c;
    return X->r0 + (1.0 - X->r0) * (m * m * m * m * m);
}

/* transport.rs:73-185 for one launched packet at (x, y) */
static void run_photon(const texp *X, orng *rng, double x, double y, orc_expose_stats *st)
{
    const orc_expose_cfg *c = X->cfg;
    st->launched++;
    const double w0 = fmax(exposure_at(X, x, y), 0.0);
    if (w0 <= 0.0) {
        st->zero_weight++;
        return;
    }
    const double th = rng_u(rng) * c->theta_max;
    const double ph = rng_u(rng) * TAU;
    const double s_t = sin(th);
    double d[3] = {s_t * cos(ph), s_t * sin(ph), cos(th)};
    normalize3(d);
    if (rng_u(rng) < fresnel(X, d[2])) {
        st->fresnel_rejected++;
        return;
    }
    double p[3] = {x, y, 0.001};
    double w = w0;
    int bounces = 0;
    while (bounces < c->max_bounces) {
        bounces++;
        st->bounces++;
        double s;
        if (X->mu_t <= 0.0) s = 1e6;
        else s = -log(fmax(1.0 - rng_u(rng), DBL_MIN)) / X->mu_t;
        const double step = fmin(s, c->D * 1.5);
        hitctx h;
        h.X = X;
        h.rng = rng;
        h.o = p;
        h.d = d;
        for (int a = 0; a < 3; a++) h.inv[a] = 1.0 / d[a];
        h.step = step;
        h.w = w;
        h.st = st;
        tcount cnt = {0, 0};
        if (X->b->nn > 0) DFS_TRAVERSE(X->b, tr_box, tr_prim, &h, &cnt);
        st->nodes_tested += cnt.nodes;
        st->prims_tested += cnt.prims;
        if (w > 0.0) {
            st->bounces_live++;
            st->nodes_live += cnt.nodes;
        }
        w = h.w;
        const double q[3] = {p[0] + d[0] * step, p[1] + d[1] * step, p[2] + d[2] * step};
        if (q[0] < 0.0 || q[0] >= c->W || q[1] < 0.0 || q[1] >= c->H) return;
        if (q[2] <= 0.0) return;
        if (q[2] >= c->D) {
            p[0] = q[0];
            p[1] = q[1];
            p[2] = c->D - 0.001;
            const double din[3] = {d[0], d[1], d[2]};
            if (!antihalation(X, &w, d, rng)) return;
            /* zero weight deposits nothing more (the reference keeps walking it) */
            if (!(w > 0.0)) {
                if (c->base_thickness > 0.0) return;
                continue;
            }
            if (c->base_thickness > 0.0) {
                /* extension: re-entry displaced by the path through the film base */
                if (!(din[2] > 1e-6) || !(-d[2] > 1e-6)) return;
                p[0] += c->base_thickness * (din[0] / din[2] + d[0] / (-d[2]));
                p[1] += c->base_thickness * (din[1] / din[2] + d[1] / (-d[2]));
                if (p[0] < 0.0 || p[0] >= c->W || p[1] < 0.0 || p[1] >= c->H) return;
            }
            continue;
        }
        p[0] = q[0];
        p[1] = q[1];
        p[2] = q[2];
        sample_hg(d, c->g, rng);
        w *= X->albedo;
        if (w < 1e-3) {
            if (rng_u(rng) < 0.1) w /= 0.1;
            else return;
        }
    }
}

static void *expose_worker(void *arg)
{
    void **a = (void **)arg;
    texp *X = (texp *)a[0];
    tstat *T = (tstat *)a[1];
    const orc_expose_cfg *c = X->cfg;
    for (;;) {
        const int64_t w = __atomic_fetch_add(&X->next, 1, __ATOMIC_RELAXED);
        if (c->sampler == 0) {
            /* one Rayon chunk: StdRng::seed_from_u64(seed + chunk * 1013904223) */
            if (w >= X->chunks) break;
            orng rng;
            rng_seq(&rng, c->seed + (uint64_t)w * 1013904223ull);
            for (int64_t i = 0; i < X->per_chunk; i++) {
                const double x = rng_u(&rng) * c->W;
                const double y = rng_u(&rng) * c->H;
                run_photon(X, &rng, x, y, &T->st);
            }
        } else {
            const int64_t p0 = w * X->grain;
            if (p0 >= X->n_total) break;
            const int64_t p1 = (p0 + X->grain < X->n_total) ? p0 + X->grain : X->n_total;
            for (int64_t p = p0; p < p1; p++) {
                orng rng;
                rng_photon(&rng, c->seed, (uint64_t)p);
                double x, y;
                if (c->ppp > 0) {
                    const int64_t pix = p / c->ppp;
                    const int64_t px = pix % c->tw, py = pix / c->tw;
                    const float ux = (float)rng_u(&rng), uy = (float)rng_u(&rng);
                    /* stratified pixel position, evaluated in float as the kernel does */
                    const float pw = (float)(c->W / (double)c->tw);
                    const float ph = (float)(c->H / (double)c->th);
                    x = (double)(((float)px + ux) * pw);
                    y = (double)(((float)py + uy) * ph);
                } else {
                    const float ux = (float)rng_u(&rng), uy = (float)rng_u(&rng);
                    x = (double)(ux * (float)c->W);
                    y = (double)(uy * (float)c->H);
                }
                run_photon(X, &rng, x, y, &T->st);
            }
        }
    }
    return NULL;
}

static int online_threads(int32_t req)
{
    if (req > 0) return req;
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

/* expose_emulsion (transport.rs:48-188).  sampler 0 mirrors the Rayon chunking
 * (chunks = 8 * threads, ceil(base / chunks) packets each, so the launched total is
 * rounded up); sampler 1 replays the GPU's photon-keyed streams. */
int orc_expose(const orc_crystals *cs, const float *exposure_map, const orc_expose_cfg *cfg,
               uint32_t *counts, orc_expose_stats *st)
{
    memset(st, 0, sizeof(*st));
    obvh b;
    bvh_build(cs, 1, &b);
    ocrys *cr = make_recs(cs);
    texp X;
    memset(&X, 0, sizeof X);
    X.b = &b;
    X.cr = cr;
    X.cfg = cfg;
    X.map = exposure_map;
    X.counts = counts;
    X.mu_t = cfg->mu_a + cfg->mu_s;
    X.albedo = X.mu_t > 0.0 ? cfg->mu_s / X.mu_t : 0.0;
    const double rr = (1.0 - cfg->n_refr) / (1.0 + cfg->n_refr);
    X.r0 = rr * rr;
    X.halation_eff = cfg->r_ahu * pow(10.0, -2.0 * cfg->od_ahu);
    const int threads = online_threads(cfg->threads);
    if (cfg->sampler == 0) {
        const int64_t base = (int64_t)llround(cfg->W * cfg->H * cfg->photons_per_area);
        if (base == 0) {
            free(cr);
            bvh_free(&b);
            return 0;
        }
        X.chunks = (int64_t)threads * 8;
        X.per_chunk = (base + X.chunks - 1) / X.chunks;
        st->chunks = X.chunks;
    } else {
        X.n_total = cfg->ppp > 0 ? (int64_t)cfg->ppp * cfg->tw * cfg->th : cfg->n_photons;
        X.grain = 4096;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    tstat *ts = (tstat *)calloc((size_t)threads, sizeof(tstat));
    void **args = (void **)malloc(sizeof(void *) * 2 * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        args[2 * t] = &X;
        args[2 * t + 1] = &ts[t];
        pthread_create(&th[t], NULL, expose_worker, &args[2 * t]);
    }
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        st->launched += ts[t].st.launched;
        st->fresnel_rejected += ts[t].st.fresnel_rejected;
        st->zero_weight += ts[t].st.zero_weight;
        st->bounces += ts[t].st.bounces;
        st->nodes_tested += ts[t].st.nodes_tested;
        st->prims_tested += ts[t].st.prims_tested;
        st->hits += ts[t].st.hits;
        st->abs_events += ts[t].st.abs_events;
        st->abs_total += ts[t].st.abs_total;
        st->bounces_live += ts[t].st.bounces_live;
        st->nodes_live += ts[t].st.nodes_live;
    }
    free(args);
    free(ts);
    free(th);
    free(cr);
    bvh_free(&b);
    return 0;
}

/* ---------------------------------------------------------------- development -- */
static double max_dim(const orc_crystals *cs, int64_t i)
{
    const double *d = cs->dims + 3 * i;
    return fmax(fmax(d[0], d[1]), d[2]);
}

/* develop_emulsion (development.rs:59-188) */
int orc_develop(const orc_crystals *cs, const uint32_t *counts, uint32_t threshold,
                double fe, double pf, uint8_t *developed, double *fe_out, int64_t *clump_id,
                orc_dev_report *rep)
{
    const int64_t n = cs->n;
    memset(rep, 0, sizeof(*rep));
    rep->total = n;
    if (n == 0) return 0;
    int64_t ndev = 0;
    for (int64_t i = 0; i < n; i++) {
        if (counts[i] >= threshold) {
            developed[i] = 1;
            fe_out[i] = fe;
            ndev++;
        } else {
            developed[i] = 0;
            fe_out[i] = 0.0;
        }
        clump_id[i] = -1;
    }
    rep->developed = ndev;
    rep->fixed = n - ndev;
    /* the radius query runs on the emulsion BVH built before development */
    orc_crystals plain = *cs;
    plain.developed = NULL;
    plain.fe = NULL;
    obvh b;
    bvh_build(&plain, 0, &b);
    double maxb = 0.0;
    for (int64_t i = 0; i < n; i++) maxb = fmax(maxb, max_dim(cs, i));
    uint8_t *visited = (uint8_t *)calloc((size_t)n, 1);
    int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *members = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t ncap = 1024;
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncap);
    int64_t clumps = 0, max_clump = 0, clumped = 0;
    for (int64_t i = 0; i < n; i++) {
        if (!developed[i] || visited[i]) continue;
        const int64_t cid = clumps++;
        visited[i] = 1;
        clump_id[i] = cid;
        int64_t qh = 0, qt = 0, nm = 0;
        queue[qt++] = i;
        members[nm++] = i;
        while (qh < qt) {
            const int64_t cur = queue[qh++];
            const double *cp = cs->pos + 3 * cur;
            const double r_cur = max_dim(cs, cur) * fe_out[cur];
            const double search = (r_cur + maxb * fe) * pf;
            int64_t k = query_radius(&b, cs, cp, search, nb, ncap);
            if (k > ncap) {
                ncap = k * 2;
                nb = (int32_t *)realloc(nb, sizeof(int32_t) * (size_t)ncap);
                k = query_radius(&b, cs, cp, search, nb, ncap);
            }
            for (int64_t t = 0; t < k; t++) {
                const int64_t j = nb[t];
                if (j == cur || visited[j] || !developed[j]) continue;
                const double r_nb = max_dim(cs, j) * fe_out[j];
                const double touch = (r_cur + r_nb) * pf;
                const double dx = cs->pos[3 * j] - cp[0], dy = cs->pos[3 * j + 1] - cp[1],
                             dz = cs->pos[3 * j + 2] - cp[2];
                if (dx * dx + dy * dy + dz * dz <= touch * touch) {
                    visited[j] = 1;
                    clump_id[j] = cid;
                    queue[qt++] = j;
                    members[nm++] = j;
                }
            }
        }
        if (nm > max_clump) max_clump = nm;
        clumped += nm;
        if (nm > 1) {
            const int64_t m = nm < 6 ? nm : 6;
            const double boost = 1.15 + 0.18 * (double)(m - 1);
            for (int64_t t = 0; t < nm; t++) fe_out[members[t]] *= boost;
        }
    }
    rep->clumps = clumps;
    rep->max_clump = max_clump;
    rep->avg_clump = clumps > 0 ? (double)clumped / (double)clumps : 0.0;
    free(nb);
    free(members);
    free(queue);
    free(visited);
    bvh_free(&b);
    return 0;
}

/* ----------------------------------------------------------------------- scan -- */
typedef struct {
    const obvh *b;
    const ocrys *cr;
    double W, H, D, sigma;
    int32_t w_px, h_px, g, ss;
    float pw, ph;
    uint64_t seed;
    float *T;
    int64_t next_row;
    int64_t nodes;
} tscan;

typedef struct {
    const ocrys *cr;
    double o[3], d[3], inv[3];
    double P;
} scanhit;

static inline int sc_box(const onode *nd, scanhit *h)
{
    return box_hit_d(nd->mn, nd->mx, h->o, h->inv);
}
static inline void sc_prim(uint32_t id, scanhit *h)
{
    const ocrys *r = &h->cr[id];
    double t0, t1;
    if (crystal_intersect_d(r->c, r->q, r->dims, r->morph, h->o, h->d, &t0, &t1))
        h->P += fmax(t1 - t0, 0.0);
}

/* Ray origin of ray (ax, ay, sx, sy) of pixel (px, py), evaluated in float: the
 * stratified jitter of transmittance.rs:110-120 extended by an ss x ss aperture. */
static void scan_origin(const tscan *S, int32_t px, int32_t py, int32_t ax, int32_t ay,
                        int32_t sx, int32_t sy, float *ox, float *oy)
{
    const uint32_t ray = (uint32_t)(((ay * S->ss + ax) * S->g + sy) * S->g + sx);
    const uint32_t pix = (uint32_t)(py * S->w_px + px);
    const uint32_t key[2] = {(uint32_t)S->seed, (uint32_t)(S->seed >> 32)};
    const uint32_t ctr[4] = {pix, ray >> 1, TAG_SCAN, 0u};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    const float jx = (float)(out[2 * (ray & 1u)] >> 8) * 0x1p-24f;
    const float jy = (float)(out[2 * (ray & 1u) + 1] >> 8) * 0x1p-24f;
    const float subx = ((float)sx + jx) / (float)S->g;
    const float suby = ((float)sy + jy) / (float)S->g;
    const float apx = ((float)ax + subx) / (float)S->ss;
    const float apy = ((float)ay + suby) / (float)S->ss;
    *ox = ((float)px + apx) * S->pw;
    *oy = ((float)py + apy) * S->ph;
}

static void *scan_worker(void *arg)
{
    tscan *S = (tscan *)arg;
    int64_t nodes = 0;
    for (;;) {
        const int64_t py = __atomic_fetch_add(&S->next_row, 1, __ATOMIC_RELAXED);
        if (py >= S->h_px) break;
        for (int32_t px = 0; px < S->w_px; px++) {
            double acc = 0.0;
            for (int32_t ay = 0; ay < S->ss; ay++)
                for (int32_t ax = 0; ax < S->ss; ax++)
                    for (int32_t sy = 0; sy < S->g; sy++)
                        for (int32_t sx = 0; sx < S->g; sx++) {
                            float ox, oy;
                            scan_origin(S, px, (int32_t)py, ax, ay, sx, sy, &ox, &oy);
                            scanhit h;
                            h.cr = S->cr;
                            h.o[0] = ox;
                            h.o[1] = oy;
                            h.o[2] = S->D + 0.1;
                            h.d[0] = 0.0;
                            h.d[1] = 0.0;
                            h.d[2] = -1.0;
                            for (int a = 0; a < 3; a++) h.inv[a] = 1.0 / h.d[a];
                            h.P = 0.0;
                            tcount cnt = {0, 0};
                            if (S->b->nn > 0) DFS_TRAVERSE(S->b, sc_box, sc_prim, &h, &cnt);
                            nodes += cnt.nodes;
                            acc += exp(-S->sigma * h.P);
                        }
            const double nrays = (double)(S->g * S->g * S->ss * S->ss);
            S->T[py * S->w_px + px] = (float)(acc / nrays);
        }
    }
    __atomic_fetch_add(&S->nodes, nodes, __ATOMIC_RELAXED);
    return NULL;
}

/* scan_emulsion (transmittance.rs:65-171) in NegativeTransmittance mode: BVH over the
 * developed crystals (expanded AABBs), g x g stratified rays from z = D + 0.1 along -z,
 * T = exp(-sigma * sum of chords), mean per pixel. */
int orc_scan(const orc_crystals *cs, double W, double H, double D, int32_t w_px,
             int32_t h_px, int32_t spp, int32_t ss, double sigma, uint64_t seed,
             int32_t threads, float *T, int64_t *rays_nodes)
{
    /* clone the developed crystals (transmittance.rs:82-87) */
    int64_t nd = 0;
    for (int64_t i = 0; i < cs->n; i++)
        if (cs->developed && cs->developed[i]) nd++;
    double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)(nd ? nd : 1));
    double *dims = (double *)malloc(sizeof(double) * 3 * (size_t)(nd ? nd : 1));
    double *quat = (double *)malloc(sizeof(double) * 4 * (size_t)(nd ? nd : 1));
    int32_t *morph = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nd ? nd : 1));
    double *fe = (double *)malloc(sizeof(double) * (size_t)(nd ? nd : 1));
    uint8_t *dev = (uint8_t *)malloc((size_t)(nd ? nd : 1));
    int64_t k = 0;
    for (int64_t i = 0; i < cs->n; i++) {
        if (!(cs->developed && cs->developed[i])) continue;
        memcpy(pos + 3 * k, cs->pos + 3 * i, 3 * sizeof(double));
        memcpy(dims + 3 * k, cs->dims + 3 * i, 3 * sizeof(double));
        memcpy(quat + 4 * k, cs->quat + 4 * i, 4 * sizeof(double));
        morph[k] = cs->morph[i];
        fe[k] = cs->fe ? cs->fe[i] : 1.0;
        dev[k] = 1;
        k++;
    }
    orc_crystals dc = {nd, pos, dims, quat, morph, NULL, dev, fe};
    obvh b;
    bvh_build(&dc, 1, &b);
    ocrys *cr = make_recs(&dc);
    tscan S;
    memset(&S, 0, sizeof S);
    S.b = &b;
    S.cr = cr;
    S.W = W;
    S.H = H;
    S.D = D;
    S.sigma = sigma;
    S.w_px = w_px;
    S.h_px = h_px;
    const int32_t sp = spp > 1 ? spp : 1;
    S.g = (int32_t)round(sqrt((double)sp));
    if (S.g < 1) S.g = 1;
    S.ss = ss > 1 ? ss : 1;
    S.pw = (float)(W / (double)w_px);
    S.ph = (float)(H / (double)h_px);
    S.seed = seed;
    S.T = T;
    const int nt = online_threads(threads);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
    for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, scan_worker, &S);
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    if (rays_nodes) *rays_nodes = S.nodes;
    free(th);
    free(cr);
    bvh_free(&b);
    free(pos);
    free(dims);
    free(quat);
    free(morph);
    free(fe);
    free(dev);
    return 0;
}

/* D_neg = -log10(clamp(T, 1e-4, 1)) + 0.05 (image_pipeline.rs:1006-1011) */
void orc_dneg(const float *T, int64_t n, float *D)
{
    for (int64_t i = 0; i < n; i++) {
        const double t = fmin(fmax((double)T[i], 1e-4), 1.0);
        D[i] = (float)(-log10(t) + 0.05);
    }
}

/* scanner OTF, separable [0.035, 0.93, 0.035] with clamp-to-edge (image_pipeline.rs:822-855) */
void orc_otf3(const float *in, int32_t w, int32_t h, float *out)
{
    const float k0 = 0.035f, k1 = 0.93f, k2 = 0.035f;
    float *tmp = (float *)malloc(sizeof(float) * (size_t)w * (size_t)h);
    for (int32_t y = 0; y < h; y++)
        for (int32_t x = 0; x < w; x++) {
            const int32_t xm = x > 0 ? x - 1 : 0, xp = x + 1 < w ? x + 1 : w - 1;
            const float *r = in + (size_t)y * w;
            tmp[(size_t)y * w + x] = r[xm] * k0 + r[x] * k1 + r[xp] * k2;
        }
    for (int32_t y = 0; y < h; y++) {
        const int32_t ym = y > 0 ? y - 1 : 0, yp = y + 1 < h ? y + 1 : h - 1;
        for (int32_t x = 0; x < w; x++)
            out[(size_t)y * w + x] = tmp[(size_t)ym * w + x] * k0 + tmp[(size_t)y * w + x] * k1 +
                                     tmp[(size_t)yp * w + x] * k2;
    }
    free(tmp);
}

/* ---------------------------------------------------------------- synthesis -- */
static double normal_sample(orng *r, double mean, double sd)
{
    /* Box-Muller, one sample per two uniforms (rand_distr's Ziggurat is not vendored) */
    const double u1 = 1.0 - rng_u(r), u2 = rng_u(r);
    return mean + sd * (sqrt(-2.0 * log(u1)) * cos(TAU * u2));
}

/* per-crystal draws of VirtualEmulsionLayer::synthesize (emulsion.rs:160-207) */
static void crystal_attrs(const orc_synth_cfg *cfg, orng *rng, double *dims, double *q, int32_t *morph,
                          double *sens)
{
    const double mu = log(cfg->mean_radius);
    double rad = exp(normal_sample(rng, mu, cfg->size_sigma));
    rad = fmin(fmax(rad, cfg->min_radius), cfg->r_min * 0.48);
    const double s = fmax(normal_sample(rng, cfg->base_sens, cfg->sens_sigma), 0.1);
    if (cfg->morph == 0 || cfg->morph == 2) {
        const double u1 = rng_u(rng), u2 = rng_u(rng), u3 = rng_u(rng);
        /* UnitQuaternion::from_euler_angles(roll, pitch, yaw) */
        const double ro = u1 * TAU * 0.5, pi_ = u2 * PI_ * 0.5, ya = u3 * TAU * 0.5;
        const double sr = sin(ro), cr = cos(ro), sp = sin(pi_), cp = cos(pi_),
                     sy = sin(ya), cy = cos(ya);
        q[0] = cr * cp * cy + sr * sp * sy;
        q[1] = sr * cp * cy - cr * sp * sy;
        q[2] = cr * sp * cy + sr * cp * sy;
        q[3] = cr * cp * sy - sr * sp * cy;
        dims[0] = dims[1] = dims[2] = rad;
    } else {
        const double ht = fmax(rad / cfg->aspect, cfg->min_half_thickness);
        const double tilt = -0.17 + 0.34 * rng_u(rng);
        double ax[3] = {-1.0 + 2.0 * rng_u(rng), -1.0 + 2.0 * rng_u(rng), 0.0};
        normalize3(ax);
        const double sh = sin(tilt * 0.5), ch = cos(tilt * 0.5);
        q[0] = ch;
        q[1] = ax[0] * sh;
        q[2] = ax[1] * sh;
        q[3] = ax[2] * sh;
        dims[0] = rad;
        dims[1] = rad;
        dims[2] = ht;
    }
    *morph = cfg->morph;
    *sens = s;
}

/* sample_poisson_3d (poisson.rs:14-132) followed by the per-crystal draws of
 * VirtualEmulsionLayer::synthesize (emulsion.rs:155-211). */
int64_t orc_synthesize(const orc_synth_cfg *cfg, uint64_t seed, int64_t cap, double *pos,
                       double *dims, double *quat, int32_t *morph, double *sens)
{
    const double W = cfg->W, H = cfg->H, D = cfg->D, r = cfg->r_min;
    if (W <= 0 || H <= 0 || D <= 0 || r <= 0) return 0;
    orng rng;
    rng_seq(&rng, seed);
    const double cell = r / sqrt(3.0);
    const int64_t gw = (int64_t)ceil(W / cell), gh = (int64_t)ceil(H / cell),
                  gd = (int64_t)ceil(D / cell);
    int64_t *grid = (int64_t *)malloc(sizeof(int64_t) * (size_t)(gw * gh * gd));
    for (int64_t i = 0; i < gw * gh * gd; i++) grid[i] = -1;
    int64_t pcap = 1024, np = 0, na = 0;
    double *pts = (double *)malloc(sizeof(double) * 3 * (size_t)pcap);
    int64_t *active = (int64_t *)malloc(sizeof(int64_t) * (size_t)pcap);
#define CELL_OF(p, gx, gy, gz)                                                      \
    do {                                                                            \
        int64_t a_ = (int64_t)floor((p)[0] / cell), b_ = (int64_t)floor((p)[1] / cell), \
                c_ = (int64_t)floor((p)[2] / cell);                                 \
        gx = a_ < 0 ? 0 : (a_ > gw - 1 ? gw - 1 : a_);                              \
        gy = b_ < 0 ? 0 : (b_ > gh - 1 ? gh - 1 : b_);                              \
        gz = c_ < 0 ? 0 : (c_ > gd - 1 ? gd - 1 : c_);                              \
    } while (0)
    {
        const double x = rng_u(&rng) * W, y = rng_u(&rng) * H, z = rng_u(&rng) * D;
        pts[0] = x;
        pts[1] = y;
        pts[2] = z;
        int64_t gx, gy, gz;
        CELL_OF(pts, gx, gy, gz);
        grid[gx + gy * gw + gz * gw * gh] = 0;
        np = 1;
        active[na++] = 0;
    }
    const double r2 = r * r;
    while (na > 0) {
        const int64_t pick = (int64_t)rng_below(&rng, (uint64_t)na);
        const double *base = pts + 3 * active[pick];
        const double bx = base[0], by = base[1], bz = base[2];
        int found = 0;
        for (int k = 0; k < cfg->k_samples; k++) {
            const double u = rng_u(&rng);
            const double ct = -1.0 + 2.0 * rng_u(&rng);
            const double stt = sqrt(fmax(1.0 - ct * ct, 0.0));
            const double phi = rng_u(&rng) * TAU;
            const double rad = cbrt(7.0 * u + 1.0) * r;
            const double cand[3] = {bx + rad * stt * cos(phi), by + rad * stt * sin(phi),
                                    bz + rad * ct};
            if (cand[0] < 0.0 || cand[0] >= W || cand[1] < 0.0 || cand[1] >= H ||
                cand[2] < 0.0 || cand[2] >= D)
                continue;
            int64_t cx, cy, cz;
            CELL_OF(cand, cx, cy, cz);
            const int64_t x0 = cx - 2 < 0 ? 0 : cx - 2, x1 = cx + 2 > gw - 1 ? gw - 1 : cx + 2;
            const int64_t y0 = cy - 2 < 0 ? 0 : cy - 2, y1 = cy + 2 > gh - 1 ? gh - 1 : cy + 2;
            const int64_t z0 = cz - 2 < 0 ? 0 : cz - 2, z1 = cz + 2 > gd - 1 ? gd - 1 : cz + 2;
            int close = 0;
            for (int64_t z = z0; z <= z1 && !close; z++)
                for (int64_t y = y0; y <= y1 && !close; y++)
                    for (int64_t x = x0; x <= x1; x++) {
                        const int64_t e = grid[x + y * gw + z * gw * gh];
                        if (e >= 0) {
                            const double dx = pts[3 * e] - cand[0], dy = pts[3 * e + 1] - cand[1],
                                         dz = pts[3 * e + 2] - cand[2];
                            if (dx * dx + dy * dy + dz * dz < r2) {
                                close = 1;
                                break;
                            }
                        }
                    }
            if (!close) {
                if (np == pcap) {
                    pcap *= 2;
                    pts = (double *)realloc(pts, sizeof(double) * 3 * (size_t)pcap);
                    active = (int64_t *)realloc(active, sizeof(int64_t) * (size_t)pcap);
                }
                grid[cx + cy * gw + cz * gw * gh] = np;
                memcpy(pts + 3 * np, cand, sizeof cand);
                active[na++] = np;
                np++;
                found = 1;
                break;
            }
        }
        if (!found) active[pick] = active[--na];
    }
#undef CELL_OF
    free(grid);
    free(active);
    if (np > cap) {
        free(pts);
        return -np;
    }
    for (int64_t i = 0; i < np; i++) {
        memcpy(pos + 3 * i, pts + 3 * i, 3 * sizeof(double));
        crystal_attrs(cfg, &rng, dims + 3 * i, quat + 4 * i, morph + i, sens + i);
    }
    free(pts);
    return np;
}

/* ------------------------------------------------- optics helpers for tests -- */
/* GelatinOptics::sample_free_path with a given uniform (optics.rs:100-107) */
double orc_free_path(double mu_t, double xi)
{
    if (mu_t <= 0.0) return 1e6;
    return -log(fmax(1.0 - xi, DBL_MIN)) / mu_t;
}

/* GelatinOptics::sample_henyey_greenstein with given uniforms (optics.rs:113-143) */
void orc_hg(const double d[3], double g, double xi1, double xi2, double out[3])
{
    double ct;
    if (fabs(g) < 1e-4) {
        ct = 2.0 * xi1 - 1.0;
    } else {
        const double term = (1.0 - g * g) / (1.0 - g + 2.0 * g * xi1);
        ct = (1.0 + g * g - term * term) / (2.0 * g);
    }
    ct = fmin(fmax(ct, -1.0), 1.0);
    const double stt = sqrt(fmax(1.0 - ct * ct, 0.0));
    const double phi = TAU * xi2;
    double w[3] = {d[0], d[1], d[2]};
    normalize3(w);
    double u[3];
    if (fabs(w[0]) > 0.1) {
        u[0] = w[2]; u[1] = 0.0; u[2] = -w[0];
    } else {
        u[0] = 0.0; u[1] = -w[2]; u[2] = w[1];
    }
    normalize3(u);
    const double v[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2],
                         w[0] * u[1] - w[1] * u[0]};
    const double a = stt * cos(phi), bb = stt * sin(phi);
    for (int k = 0; k < 3; k++) out[k] = (u[k] * a + v[k] * bb) + w[k] * ct;
    normalize3(out);
}

/* GelatinOptics::fresnel_air_gelatin (optics.rs:147-154) */
double orc_fresnel(double n, double cos_i)
{
    texp X;
    memset(&X, 0, sizeof X);
    const double r = (1.0 - n) / (1.0 + n);
    X.r0 = r * r;
    return fresnel(&X, cos_i);
}

/* handle_antihalation_reflection (optics.rs:160-196) with a sequential stream seeded
 * by seed; returns alive and updates energy and direction. */
int orc_antihalation(double r_ahu, double od, uint64_t seed, double *energy, double d[3])
{
    texp X;
    memset(&X, 0, sizeof X);
    X.halation_eff = r_ahu * pow(10.0, -2.0 * od);
    orng rng;
    rng_seq(&rng, seed);
    return antihalation(&X, energy, d, &rng);
}

/* Chord of crystal record r with semi-axes scaled by s and d chord / d s.  The chord is
 * crystal_intersect on the scaled record; the derivative: box -- the exit plane (and the
 * entry plane when the ray starts outside) moves by dims_k / |d_k| per unit s; ellipsoid
 * -- in the unit frame (P = l / dims, Q = d / dims) the roots are
 * (-P.Q -+ sqrt((P.Q)^2 - Q.Q (P.P - s^2))) / Q.Q, each moving by s / sqrt(...). */
static int chord_scaled_d(const ocrys *r, double s, const double o[3], const double d[3],
                          double *chord, double *dchord)
{
    const double hs[3] = {r->dims[0] * s, r->dims[1] * s, r->dims[2] * s};
    double t0, t1;
    if (!crystal_intersect_d(r->c, r->q, hs, r->morph, o, d, &t0, &t1)) return 0;
    *chord = fmax(t1 - t0, 0.0);
    double rel[3] = {o[0] - r->c[0], o[1] - r->c[1], o[2] - r->c[2]}, l[3], e[3];
    qinv_rotate_d(r->q, rel, l);
    qinv_rotate_d(r->q, d, e);
    if (r->morph == 0) {
        double tmin = -INFINITY, tmax = INFINITY, gmin = 0.0, gmax = 0.0;
        for (int k = 0; k < 3; k++) {
            const double inv = 1.0 / e[k];
            const double a1 = (-hs[k] - l[k]) * inv, a2 = (hs[k] - l[k]) * inv;
            const double lo = fmin(a1, a2), hi = fmax(a1, a2);
            const double g = r->dims[k] * fabs(inv);
            if (lo > tmin) { tmin = lo; gmin = g; }
            if (hi < tmax) { tmax = hi; gmax = g; }
        }
        *dchord = gmax + (tmin > 0.0 ? gmin : 0.0);
        if (!(*dchord < INFINITY)) *dchord = 0.0;
        return 1;
    }
    double P[3], Q[3];
    for (int k = 0; k < 3; k++) {
        P[k] = l[k] / r->dims[k];
        Q[k] = e[k] / r->dims[k];
    }
    const double QQ = Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2];
    const double PQ = P[0] * Q[0] + P[1] * Q[1] + P[2] * Q[2];
    const double PP = P[0] * P[0] + P[1] * P[1] + P[2] * P[2];
    const double rt = sqrt(fmax(PQ * PQ - QQ * (PP - s * s), 0.0));
    const double enter = (-PQ - rt) / QQ;
    *dchord = rt > 0.0 ? (enter > 0.0 ? 2.0 : 1.0) * s / rt : 0.0;
    return 1;
}

int orc_chord_scaled(const orc_crystals *cs, int64_t i, double s, const double o[3],
                     const double d[3], double out[2])
{
    ocrys r;
    make_rec(cs, i, &r);
    return chord_scaled_d(&r, s, o, d, &out[0], &out[1]);
}

/* =====================================================================================
 * Expected (differentiable) model -- no reference counterpart.  FP64 twin of the
 * library's path replay: the packets of the photon-keyed sampler (sampler 1, ppp > 0)
 * without the stochastic-rounding and roulette draws; every hit deposits its mean
 * latent count 5 * E * w * pa * sens (transport.rs:117-134); a packet ends when it
 * leaves the slab, after max_bounces, or once its relative weight w and the weight's
 * reflectance tangent are both below the cutoff.  Gradients: forward tangents of w for
 * mu_abs and the antihalation reflectance R, the Henyey-Greenstein score for g.
 * ===================================================================================== */
typedef struct {
    const obvh *b;
    const ocrys *cr;
    const orc_expected_cfg *e;
    const float *map;
    const double *gdose;    /* NULL: dose pass */
    double mu_t, albedo, r0, hal_unit, hal_eff, cut;
    int64_t n_total, grain, next;
    int64_t n;              /* crystals */
} trep;

typedef struct {
    double *dose;           /* dose pass: n */
    double *gmap;           /* gradient pass: tw * th */
    double gmu, gg, gr;
    int64_t bounces, events;
} trep_acc;

typedef struct {
    const trep *X;
    trep_acc *A;
    const double *o, *d;
    double inv[3];
    double step, E;
    double w, w_mu, w_r, score;
    double a_e, a_mu, a_g, a_r;
} rephit;

static inline int rp_box(const onode *nd, rephit *h)
{
    return box_hit_d(nd->mn, nd->mx, h->o, h->inv);
}

static inline void rp_prim(uint32_t id, rephit *h)
{
    const ocrys *r = &h->X->cr[id];
    double t0, t1;
    if (!crystal_intersect_d(r->c, r->q, r->dims, r->morph, h->o, h->d, &t0, &t1)) return;
    if (!(t0 < h->step)) return;
    const double L = fmax(fmin(t1, h->step) - t0, 0.0);
    if (!(L > 0.0)) return;
    h->A->events++;
    const double pa = 1.0 - exp(-h->X->e->x.mu_abs * L);
    if (h->X->gdose) {
        const double base = 5.0 * r->sens * h->X->gdose[id];
        const double dep = h->w * pa;
        h->a_e += base * dep;
        h->a_mu += base * h->E * (h->w_mu * pa + h->w * L * (1.0 - pa));
        h->a_r += base * h->E * (h->w_r * pa);
        h->a_g += base * h->E * dep * h->score;
    } else {
        h->A->dose[id] += h->E * h->w * pa * r->sens * 5.0;
    }
    const double m = 1.0 - 0.9 * pa, dm = -0.9 * L * (1.0 - pa);
    h->w_mu = h->w_mu * m + h->w * dm;
    h->w_r *= m;
    h->w *= m;
}

/* d/dg log p_g(mu) of the Henyey-Greenstein density of mu = cos(theta) */
double orc_hg_score(double g, double mu)
{
    return -2.0 * g / (1.0 - g * g) + 3.0 * (mu - g) / (1.0 + g * g - 2.0 * g * mu);
}

static void replay_packet(const trep *X, trep_acc *A, int64_t p)
{
    const orc_expose_cfg *c = &X->e->x;
    orng rng;
    rng_photon(&rng, c->seed, (uint64_t)p);
    const int64_t pix = p / c->ppp;
    const int64_t px = pix % c->tw, py = pix / c->tw;
    const float ux = (float)rng_u(&rng), uy = (float)rng_u(&rng);
    const float pw = (float)(c->W / (double)c->tw), ph = (float)(c->H / (double)c->th);
    const double x = (double)(((float)px + ux) * pw), y = (double)(((float)py + uy) * ph);
    /* nearest-pixel lookup (image_pipeline.rs:949-961) */
    const double u = fmin(fmax(x / c->W, 0.0), 1.0), v = fmin(fmax(y / c->H, 0.0), 1.0);
    int64_t ex = (int64_t)round(u * (double)(c->tw - 1)), ey = (int64_t)round(v * (double)(c->th - 1));
    if (ex > c->tw - 1) ex = c->tw - 1;
    if (ey > c->th - 1) ey = c->th - 1;
    const int64_t e_idx = ey * c->tw + ex;
    const double E = fmax((double)X->map[e_idx], 0.0);
    if (!(E > 0.0)) return;
    const double th = rng_u(&rng) * c->theta_max;
    const double phi = rng_u(&rng) * TAU;
    const double s_t = sin(th);
    double d[3] = {s_t * cos(phi), s_t * sin(phi), cos(th)};
    normalize3(d);
    {
        const double cz = fmin(fmax(fabs(d[2]), 0.0), 1.0), m = 1.0 - cz;
        if (rng_u(&rng) < X->r0 + (1.0 - X->r0) * (m * m * m * m * m)) return;
    }
    double o[3] = {x, y, 0.001};
    rephit h;
    memset(&h, 0, sizeof h);
    h.X = X;
    h.A = A;
    h.o = o;
    h.d = d;
    h.E = E;
    h.w = 1.0;
    for (int b = 0; b < c->max_bounces; b++) {
        A->bounces++;
        const double s = X->mu_t <= 0.0 ? 1e6 : -log(fmax(1.0 - rng_u(&rng), DBL_MIN)) / X->mu_t;
        h.step = fmin(s, c->D * 1.5);
        for (int a = 0; a < 3; a++) h.inv[a] = 1.0 / d[a];
        tcount cnt = {0, 0};
        if (X->b->nn > 0) DFS_TRAVERSE(X->b, rp_box, rp_prim, &h, &cnt);
        const double q[3] = {o[0] + d[0] * h.step, o[1] + d[1] * h.step, o[2] + d[2] * h.step};
        if (q[0] < 0.0 || q[0] >= c->W || q[1] < 0.0 || q[1] >= c->H) break;
        if (q[2] <= 0.0) break;
        if (q[2] >= c->D) {
            o[0] = q[0];
            o[1] = q[1];
            o[2] = c->D - 0.001;
            const double din[3] = {d[0], d[1], d[2]};
            h.w_r = h.w_r * X->hal_eff + h.w * X->hal_unit;
            h.w_mu *= X->hal_eff;
            h.w *= X->hal_eff;
            if (rng_u(&rng) < 0.4) {
                const double u1 = rng_u(&rng), u2 = rng_u(&rng);
                const double r = sqrt(u1), ph2 = TAU * u2;
                d[0] = r * cos(ph2);
                d[1] = r * sin(ph2);
                d[2] = -sqrt(fmax(1.0 - u1, 0.0));
                normalize3(d);
            } else {
                d[2] = -fabs(d[2]);
                normalize3(d);
            }
            if (h.w < X->cut && h.w_r < X->cut) break;
            if (c->base_thickness > 0.0) {
                if (!(din[2] > 1e-6) || !(-d[2] > 1e-6)) break;
                o[0] += c->base_thickness * (din[0] / din[2] + d[0] / (-d[2]));
                o[1] += c->base_thickness * (din[1] / din[2] + d[1] / (-d[2]));
                if (o[0] < 0.0 || o[0] >= c->W || o[1] < 0.0 || o[1] >= c->H) break;
            }
            continue;
        }
        o[0] = q[0];
        o[1] = q[1];
        o[2] = q[2];
        const double ct = sample_hg(d, c->g, &rng);
        h.score += orc_hg_score(c->g, ct);
        h.w *= X->albedo;
        h.w_mu *= X->albedo;
        h.w_r *= X->albedo;
        if (h.w < X->cut && h.w_r < X->cut) break;
    }
    if (X->gdose) {
        A->gmap[e_idx] += h.a_e;
        A->gmu += h.a_mu;
        A->gg += h.a_g;
        A->gr += h.a_r;
    }
}

static void *replay_worker(void *arg)
{
    void **a = (void **)arg;
    trep *X = (trep *)a[0];
    trep_acc *A = (trep_acc *)a[1];
    for (;;) {
        const int64_t p0 = __atomic_fetch_add(&X->next, X->grain, __ATOMIC_RELAXED);
        if (p0 >= X->n_total) break;
        const int64_t p1 = p0 + X->grain < X->n_total ? p0 + X->grain : X->n_total;
        for (int64_t p = p0; p < p1; p++) replay_packet(X, A, p);
    }
    return NULL;
}

/* one replay pass; gdose == NULL: dose (n), else gmap (tw * th) and grads[3] */
static int replay_pass(const orc_crystals *cs, const float *map, const orc_expected_cfg *e,
                       const double *gdose, double *dose, double *gmap, double *grads,
                       int64_t *bounces_events)
{
    const orc_expose_cfg *c = &e->x;
    if (c->ppp <= 0 || c->tw <= 0 || c->th <= 0) return -1;
    orc_crystals plain = *cs;
    plain.developed = NULL;
    plain.fe = NULL;
    obvh b;
    bvh_build(&plain, 1, &b);
    ocrys *cr = make_recs(&plain);
    trep X;
    memset(&X, 0, sizeof X);
    X.b = &b;
    X.cr = cr;
    X.e = e;
    X.map = map;
    X.gdose = gdose;
    X.n = cs->n;
    X.mu_t = c->mu_a + c->mu_s;
    X.albedo = X.mu_t > 0.0 ? c->mu_s / X.mu_t : 0.0;
    const double rr = (1.0 - c->n_refr) / (1.0 + c->n_refr);
    X.r0 = rr * rr;
    X.hal_unit = pow(10.0, -2.0 * c->od_ahu);
    X.hal_eff = c->r_ahu * X.hal_unit;
    X.cut = e->cut > 0.0 ? e->cut : 1e-6;
    X.n_total = (int64_t)c->ppp * c->tw * c->th;
    X.grain = 2048;
    const int nt = online_threads(c->threads);
    const int64_t nmap = (int64_t)c->tw * c->th;
    trep_acc *acc = (trep_acc *)calloc((size_t)nt, sizeof(trep_acc));
    for (int t = 0; t < nt; t++) {
        if (gdose) acc[t].gmap = (double *)calloc((size_t)nmap, sizeof(double));
        else acc[t].dose = (double *)calloc((size_t)(cs->n > 0 ? cs->n : 1), sizeof(double));
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
    void **args = (void **)malloc(sizeof(void *) * 2 * (size_t)nt);
    for (int t = 0; t < nt; t++) {
        args[2 * t] = &X;
        args[2 * t + 1] = &acc[t];
        pthread_create(&th[t], NULL, replay_worker, &args[2 * t]);
    }
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    if (gdose) {
        memset(gmap, 0, sizeof(double) * (size_t)nmap);
        grads[0] = grads[1] = grads[2] = 0.0;
    } else {
        memset(dose, 0, sizeof(double) * (size_t)cs->n);
    }
    if (bounces_events) bounces_events[0] = bounces_events[1] = 0;
    for (int t = 0; t < nt; t++) {
        if (gdose) {
            for (int64_t k = 0; k < nmap; k++) gmap[k] += acc[t].gmap[k];
            grads[0] += acc[t].gmu;
            grads[1] += acc[t].gg;
            grads[2] += acc[t].gr;
            free(acc[t].gmap);
        } else {
            for (int64_t k = 0; k < cs->n; k++) dose[k] += acc[t].dose[k];
            free(acc[t].dose);
        }
        if (bounces_events) {
            bounces_events[0] += acc[t].bounces;
            bounces_events[1] += acc[t].events;
        }
    }
    free(args);
    free(th);
    free(acc);
    free(cr);
    bvh_free(&b);
    return 0;
}

int orc_expected_dose(const orc_crystals *cs, const float *map, const orc_expected_cfg *e,
                      double *dose, int64_t *bounces_events)
{
    return replay_pass(cs, map, e, NULL, dose, NULL, NULL, bounces_events);
}

int orc_replay_grad(const orc_crystals *cs, const float *map, const orc_expected_cfg *e,
                    const double *gdose, double *gmap, double grads[3])
{
    return replay_pass(cs, map, e, gdose, NULL, gmap, grads, NULL);
}

/* P[Poisson(lam) >= thr] and its derivative (the Poisson pmf at thr - 1) */
void orc_poisson_tail(const double *lam, int64_t n, uint32_t thr, double *p, double *dp)
{
    for (int64_t i = 0; i < n; i++) {
        const double l = lam[i];
        if (thr == 0) {
            p[i] = 1.0;
            dp[i] = 0.0;
            continue;
        }
        if (!(l > 0.0)) {
            p[i] = 0.0;
            dp[i] = thr == 1 ? 1.0 : 0.0;
            continue;
        }
        /* terms l^k e^-l / k! in log space: the tail sum below thr, else 1 - head */
        if (l < (double)thr) {
            double tail = 0.0;
            for (uint32_t k = thr; k < thr + 4000u; k++) {
                const double t = exp((double)k * log(l) - l - lgamma((double)k + 1.0));
                tail += t;
                if (t < 1e-18 * tail) break;
            }
            p[i] = tail;
        } else {
            double head = 0.0;
            for (uint32_t k = 0; k < thr; k++) head += exp((double)k * log(l) - l - lgamma((double)k + 1.0));
            p[i] = 1.0 - head;
        }
        dp[i] = exp((double)(thr - 1) * log(l) - l - lgamma((double)thr));
    }
}

/* ----------------------------------------------------------- expected scan ------ */
typedef struct {
    tscan S;                 /* geometry, rays, seed */
    const double *prob;      /* per crystal */
    const double *boost;     /* expected clump boost per crystal */
    const double *gT;        /* adjoint: dL/dT per pixel, or NULL */
    double *T;               /* forward output */
    double *gprob;           /* adjoint: per thread n, folded after */
    int64_t n;
} texpscan;

typedef struct {
    const ocrys *cr;
    const double *prob, *boost;
    double o[3], d[3], inv[3];
    double sigma, prod, coef;
    double *gprob, *gboost;  /* NULL: product pass */
} exphit;

static inline int ex_box(const onode *nd, exphit *h)
{
    return box_hit_d(nd->mn, nd->mx, h->o, h->inv);
}
static inline void ex_prim(uint32_t id, exphit *h)
{
    double c, dc;
    if (!chord_scaled_d(&h->cr[id], h->boost[id], h->o, h->d, &c, &dc)) return;
    const double ex = exp(-h->sigma * c);
    const double a = 1.0 - ex;
    const double pr = h->prob[id];
    const double f = 1.0 - pr * a;
    if (h->gprob) {
        if (f > 0.0) {
            h->gprob[id] += -a * h->coef / f;
            h->gboost[id] += -pr * h->sigma * ex * dc * h->coef / f;
        }
    } else {
        h->prod *= f;
    }
}

typedef struct {
    texpscan *X;
    double *gprob, *gboost;
} expscan_arg;

static void *expscan_worker(void *arg)
{
    expscan_arg *a = (expscan_arg *)arg;
    texpscan *X = a->X;
    tscan *S = &X->S;
    const double nrays = (double)(S->g * S->g * S->ss * S->ss);
    for (;;) {
        const int64_t py = __atomic_fetch_add(&S->next_row, 1, __ATOMIC_RELAXED);
        if (py >= S->h_px) break;
        for (int32_t px = 0; px < S->w_px; px++) {
            double gT = X->gT ? X->gT[py * S->w_px + px] : 0.0;
            if (X->gT && gT == 0.0) continue;
            double acc = 0.0;
            for (int32_t ay = 0; ay < S->ss; ay++)
                for (int32_t ax = 0; ax < S->ss; ax++)
                    for (int32_t sy = 0; sy < S->g; sy++)
                        for (int32_t sx = 0; sx < S->g; sx++) {
                            float ox, oy;
                            scan_origin(S, px, (int32_t)py, ax, ay, sx, sy, &ox, &oy);
                            exphit h;
                            memset(&h, 0, sizeof h);
                            h.cr = S->cr;
                            h.prob = X->prob;
                            h.boost = X->boost;
                            h.o[0] = ox;
                            h.o[1] = oy;
                            h.o[2] = S->D + 0.1;
                            h.d[2] = -1.0;
                            for (int k = 0; k < 3; k++) h.inv[k] = 1.0 / h.d[k];
                            h.sigma = S->sigma;
                            h.prod = 1.0;
                            tcount cnt = {0, 0};
                            if (S->b->nn > 0) DFS_TRAVERSE(S->b, ex_box, ex_prim, &h, &cnt);
                            if (X->gT) {
                                h.coef = gT / nrays * h.prod;
                                h.gprob = a->gprob;
                                h.gboost = a->gboost;
                                if (S->b->nn > 0) DFS_TRAVERSE(S->b, ex_box, ex_prim, &h, &cnt);
                            }
                            acc += h.prod;
                        }
            if (!X->gT) X->T[py * S->w_px + px] = acc / nrays;
        }
    }
    return NULL;
}

/* Expected scan over every crystal with semi-axes times fe * boost (boxes for the largest
 * boost, fe * 2.05); gT == NULL: T per pixel (mean ray product); else gprob and gboost =
 * sum over rays of (gT / rays) * dT_ray / d(prob, boost). */
int orc_expected_scan(const orc_crystals *cs, const double *prob, const double *boost,
                      const orc_expected_cfg *e, const double *gT, double *T, double *gprob,
                      double *gboost)
{
    const int64_t n = cs->n;
    uint8_t *dev = (uint8_t *)malloc((size_t)(n > 0 ? n : 1));
    double *fe = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *fbox = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) {
        dev[i] = 1;
        fe[i] = e->fe;
        fbox[i] = e->fe * 2.05;
    }
    orc_crystals dc = *cs, db = *cs;
    dc.developed = dev;
    dc.fe = fe;
    db.developed = dev;
    db.fe = fbox;
    obvh b;
    bvh_build(&db, 1, &b);
    ocrys *cr = make_recs(&dc);
    texpscan X;
    memset(&X, 0, sizeof X);
    tscan *S = &X.S;
    S->b = &b;
    S->cr = cr;
    S->W = e->x.W;
    S->H = e->x.H;
    S->D = e->x.D;
    S->sigma = e->sigma;
    S->w_px = e->w_px;
    S->h_px = e->h_px;
    const int32_t sp = e->spp > 1 ? e->spp : 1;
    S->g = (int32_t)round(sqrt((double)sp));
    if (S->g < 1) S->g = 1;
    S->ss = e->ss > 1 ? e->ss : 1;
    S->pw = (float)(e->x.W / (double)e->w_px);
    S->ph = (float)(e->x.H / (double)e->h_px);
    S->seed = e->scan_seed;
    X.prob = prob;
    X.boost = boost;
    X.gT = gT;
    X.T = T;
    X.n = n;
    const int nt = online_threads(e->x.threads);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
    expscan_arg *args = (expscan_arg *)calloc((size_t)nt, sizeof(expscan_arg));
    for (int t = 0; t < nt; t++) {
        args[t].X = &X;
        if (gT) {
            args[t].gprob = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
            args[t].gboost = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
        }
        pthread_create(&th[t], NULL, expscan_worker, &args[t]);
    }
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    if (gT) {
        memset(gprob, 0, sizeof(double) * (size_t)n);
        memset(gboost, 0, sizeof(double) * (size_t)n);
        for (int t = 0; t < nt; t++) {
            for (int64_t i = 0; i < n; i++) {
                gprob[i] += args[t].gprob[i];
                gboost[i] += args[t].gboost[i];
            }
            free(args[t].gprob);
            free(args[t].gboost);
        }
    }
    free(args);
    free(th);
    free(cr);
    bvh_free(&b);
    free(fbox);
    free(fe);
    free(dev);
    return 0;
}

/* Touch graph of development.rs:132-155 at unboosted expansion fe: j ~ i when
 * |p_i - p_j| <= (m_i + m_j) fe pf, m the largest semi-axis.  CSR (ptr n + 1, idx);
 * returns the number of entries, or minus it when cap is too small. */
int64_t orc_touch_graph(const orc_crystals *cs, double fe, double pf, int64_t *ptr, int32_t *idx,
                        int64_t cap)
{
    const int64_t n = cs->n;
    double maxm = 0.0;
    for (int64_t i = 0; i < n; i++) maxm = fmax(maxm, max_dim(cs, i));
    const double r = 2.0 * maxm * fe * pf;
    orc_crystals plain = *cs;
    plain.developed = NULL;
    plain.fe = NULL;
    obvh b;
    bvh_build(&plain, 0, &b);
    int64_t ncap = 1024, k = 0;
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncap);
    ptr[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        const double *p = cs->pos + 3 * i;
        int64_t m = query_radius(&b, cs, p, r, nb, ncap);
        if (m > ncap) {
            ncap = m * 2;
            nb = (int32_t *)realloc(nb, sizeof(int32_t) * (size_t)ncap);
            m = query_radius(&b, cs, p, r, nb, ncap);
        }
        /* sorted neighbour list */
        for (int64_t a = 1; a < m; a++)
            for (int64_t c = a; c > 0 && nb[c - 1] > nb[c]; c--) {
                const int32_t t = nb[c];
                nb[c] = nb[c - 1];
                nb[c - 1] = t;
            }
        for (int64_t t = 0; t < m; t++) {
            const int64_t j = nb[t];
            if (j == i) continue;
            const double touch = (max_dim(cs, i) + max_dim(cs, j)) * fe * pf;
            const double dx = cs->pos[3 * j] - p[0], dy = cs->pos[3 * j + 1] - p[1],
                         dz = cs->pos[3 * j + 2] - p[2];
            if (dx * dx + dy * dy + dz * dz <= touch * touch) {
                if (k < cap) idx[k] = (int32_t)j;
                k++;
            }
        }
        ptr[i + 1] = k;
    }
    free(nb);
    bvh_free(&b);
    return k <= cap ? k : -k;
}

/* Device synthesis twin (no reference counterpart): the library's parallel dart throwing
 * evaluated sequentially in float with the kernel's operation order (cells of one phase
 * do not interact, so their order is free), then per-crystal streams for the
 * attributes.  Returns the number of crystals, or minus it when cap is too small. */
int64_t orc_darts(const orc_synth_cfg *cfg, uint64_t seed, int64_t cap, double *pos, double *dims,
                  double *quat, int32_t *morph, double *sens)
{
    const double W = cfg->W, H = cfg->H, D = cfg->D, r = cfg->r_min;
    if (W <= 0 || H <= 0 || D <= 0 || r <= 0) return 0;
    const float cell = (float)(r / sqrt(3.0));
    const uint32_t gx = (uint32_t)ceil(W / (double)cell), gy = (uint32_t)ceil(H / (double)cell),
                   gz = (uint32_t)ceil(D / (double)cell);
    const float r2 = (float)(r * r), Wf = (float)W, Hf = (float)H, Df = (float)D;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const size_t nc = (size_t)gx * gy * gz;
    float *cx = (float *)calloc(nc * 4, sizeof(float));
    const int rounds = cfg->k_samples > 0 ? cfg->k_samples : 30;
    for (int rd = 0; rd < rounds; rd++)
        for (uint32_t k0 = 0; k0 < 3; k0++)
            for (uint32_t j0 = 0; j0 < 3; j0++)
                for (uint32_t i0 = 0; i0 < 3; i0++)
                    for (uint32_t k = k0; k < gz; k += 3)
                        for (uint32_t j = j0; j < gy; j += 3)
                            for (uint32_t i = i0; i < gx; i += 3) {
                                const uint32_t c = i + gx * (j + gy * k);
                                if (cx[4 * (size_t)c + 3] != 0.0f) continue;
                                const uint32_t ctr[4] = {c, (uint32_t)rd, TAG_DART, 0u};
                                uint32_t w[4];
                                orc_philox4x32_10(ctr, key, w);
                                const float x = ((float)i + (float)(w[0] >> 8) * 0x1p-24f) * cell;
                                const float y = ((float)j + (float)(w[1] >> 8) * 0x1p-24f) * cell;
                                const float z = ((float)k + (float)(w[2] >> 8) * 0x1p-24f) * cell;
                                if (!(x < Wf && y < Hf && z < Df)) continue;
                                int ok = 1;
                                for (int dz = -2; dz <= 2 && ok; dz++) {
                                    const int kk = (int)k + dz;
                                    if (kk < 0 || kk >= (int)gz) continue;
                                    for (int dy = -2; dy <= 2 && ok; dy++) {
                                        const int jj = (int)j + dy;
                                        if (jj < 0 || jj >= (int)gy) continue;
                                        for (int dx = -2; dx <= 2; dx++) {
                                            const int ii = (int)i + dx;
                                            if (ii < 0 || ii >= (int)gx) continue;
                                            const float *q = cx + 4 * ((size_t)ii + (size_t)gx * ((size_t)jj + (size_t)gy * (size_t)kk));
                                            if (q[3] == 0.0f) continue;
                                            const float ex = q[0] - x, ey = q[1] - y, ez = q[2] - z;
                                            if ((ex * ex + ey * ey) + ez * ez < r2) {
                                                ok = 0;
                                                break;
                                            }
                                        }
                                    }
                                }
                                if (!ok) continue;
                                cx[4 * (size_t)c] = x;
                                cx[4 * (size_t)c + 1] = y;
                                cx[4 * (size_t)c + 2] = z;
                                cx[4 * (size_t)c + 3] = 1.0f;
                            }
    int64_t np = 0;
    for (size_t c = 0; c < nc; c++) np += cx[4 * c + 3] != 0.0f;
    if (np > cap) {
        free(cx);
        return -np;
    }
    int64_t p = 0;
    for (size_t c = 0; c < nc; c++) {
        if (cx[4 * c + 3] == 0.0f) continue;
        for (int a = 0; a < 3; a++) pos[3 * p + a] = (double)cx[4 * c + a];
        orng rng;
        rng_point(&rng, seed, (uint64_t)p);
        crystal_attrs(cfg, &rng, dims + 3 * p, quat + 4 * p, morph + p, sens + p);
        p++;
    }
    free(cx);
    return np;
}
