/*
 * ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's CPU
 * baseline).  Plain-C restatement of the reference's full-frame black-and-white
 * compositing (GuyNguyen/fgrain src/image_pipeline.rs):
 *   orc_film_luma       convert_to_panchromatic_luma          :259-296
 *   orc_film_plan       grain scale, tile step and overlap    :346-385
 *   orc_film_optics     apply_camera_and_gelatin_optics       :133-252
 *   orc_film_composite  Hann accumulation, low-frequency separation, sensitometric
 *                       compositing, scanner OTF, 8-bit output :593-891
 * Every float operation is single precision in the reference's order (the reference
 * computes these buffers in f32); the library is built with -ffp-contract=off, and
 * expf/cosf/sinf/powf are glibc's, which is what Rust's f32 methods call on Linux.
 *
 * The reference's mean and standard deviation of the high-pass field are sequential
 * f32 folds (`.sum::<f32>()`, :719-724); cfg->fold = 0 restates that, cfg->fold = 1 sums
 * the same f32 terms in double precision (what the MI355X kernels do).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const float PI_F = 3.14159265358979323846f;

/* convert_to_panchromatic_luma (image_pipeline.rs:259-296): gray input is returned
 * unchanged; RGB uses the preset's spectral weights, clamp, round half away from zero */
void orc_film_luma(const uint8_t *rgb, int64_t n, float rw, float gw, float bw, uint8_t *out)
{
    for (int64_t i = 0; i < n; i++) {
        const float r = (float)rgb[3 * i], g = (float)rgb[3 * i + 1], b = (float)rgb[3 * i + 2];
        float l = r * rw + g * gw + b * bw;
        l = l < 0.0f ? 0.0f : (l > 255.0f ? 255.0f : l);
        out[i] = (uint8_t)roundf(l);
    }
}

/* render_film_black_and_white_tiled's derived sizes (image_pipeline.rs:346-385) */
int orc_film_plan(double grain_size_px, double grain_strength, double scale, int32_t auto_scale,
                  uint32_t reference_res, uint32_t tile_size, uint32_t out_w, uint32_t out_h,
                  orc_film_plan_out *p)
{
    const double user_scale = scale > 0.0 ? scale : 1.0;
    double res_scale = 1.0, selwyn = 1.0;
    if (auto_scale) {
        const double ref = fmax((double)reference_res, 512.0);
        const double long_edge = fmax((double)(out_w > out_h ? out_w : out_h), 1.0);
        double ratio = long_edge / ref;
        ratio = ratio < 0.05 ? 0.05 : (ratio > 5.0 ? 5.0 : ratio);
        res_scale = ratio;
        selwyn = sqrt(ratio);
    }
    const double total = user_scale * res_scale;
    double g = grain_size_px * total;
    g = g < 0.05 ? 0.05 : (g > 15.0 ? 15.0 : g);
    p->grain_size_px = g;
    p->strength = fmax(grain_strength * sqrt(user_scale) * selwyn, 0.0);
    if (tile_size < 48) return -1; /* u32::clamp(48, tile_size) panics in the reference */
    uint32_t step = (uint32_t)llround(384.0 * g);
    step = step < 48 ? 48 : (step > tile_size ? tile_size : step);
    p->tile_step = step;
    uint32_t pad = step / 10;
    p->pad = pad < 12 ? 12 : (pad > 48 ? 48 : pad);
    p->tiles_x = (out_w + step - 1) / step;
    p->tiles_y = (out_h + step - 1) / step;
    return 0;
}

/* one separable pass with taps restricted to [lo, hi) and renormalised by the weights
 * used (image_pipeline.rs:172-186 and :661-690) */
static float tap_sum(const float *src, int32_t stride, int32_t pos, int32_t lo, int32_t hi,
                     const float *k, int32_t rad)
{
    float sum = 0.0f, wsum = 0.0f;
    for (int32_t j = 0; j <= 2 * rad; j++) {
        const int32_t tap = pos + j - rad;
        if (tap >= lo && tap < hi) {
            sum += src[(int64_t)tap * stride] * k[j];
            wsum += k[j];
        }
    }
    return sum / fmaxf(wsum, 1e-6f);
}

/* apply_camera_and_gelatin_optics (image_pipeline.rs:133-252) */
void orc_film_optics(const uint8_t *src, int32_t w, int32_t h, float *h_eff, int32_t margins[4])
{
    const int64_t n = (int64_t)w * h;
    /* gate: rows/columns that are entirely zero are cut (:137-153) */
    int32_t top = 0, bottom = h, left = 0, right = w;
    for (; top < h; top++) {
        int any = 0;
        for (int32_t x = 0; x < w && !any; x++) any = src[(int64_t)top * w + x] != 0;
        if (any) break;
    }
    for (; bottom > top; bottom--) {
        int any = 0;
        for (int32_t x = 0; x < w && !any; x++) any = src[(int64_t)(bottom - 1) * w + x] != 0;
        if (any) break;
    }
    for (; left < w; left++) {
        int any = 0;
        for (int32_t y = top; y < bottom && !any; y++) any = src[(int64_t)y * w + left] != 0;
        if (any) break;
    }
    for (; right > left; right--) {
        int any = 0;
        for (int32_t y = top; y < bottom && !any; y++) any = src[(int64_t)y * w + right - 1] != 0;
        if (any) break;
    }
    if (margins) {
        margins[0] = top;
        margins[1] = bottom;
        margins[2] = left;
        margins[3] = right;
    }
    float *lin = (float *)calloc((size_t)n, sizeof(float));
    float *tn = (float *)calloc((size_t)n, sizeof(float));
    float *on = (float *)calloc((size_t)n, sizeof(float));
    float *tw = (float *)calloc((size_t)n, sizeof(float));
    float *ow = (float *)calloc((size_t)n, sizeof(float));
    for (int32_t y = top; y < bottom; y++)
        for (int32_t x = left; x < right; x++) lin[(int64_t)y * w + x] = (float)src[(int64_t)y * w + x] / 255.0f;
    static const float kn[3] = {0.15f, 0.70f, 0.15f};
    static const float kw[7] = {0.03f, 0.12f, 0.20f, 0.30f, 0.20f, 0.12f, 0.03f};
    for (int32_t y = top; y < bottom; y++)
        for (int32_t x = left; x < right; x++)
            tn[(int64_t)y * w + x] = tap_sum(lin + (int64_t)y * w, 1, x, left, right, kn, 1);
    for (int32_t y = top; y < bottom; y++)
        for (int32_t x = left; x < right; x++)
            on[(int64_t)y * w + x] = tap_sum(tn + x, w, y, top, bottom, kn, 1);
    for (int32_t y = top; y < bottom; y++)
        for (int32_t x = left; x < right; x++)
            tw[(int64_t)y * w + x] = tap_sum(lin + (int64_t)y * w, 1, x, left, right, kw, 3);
    for (int32_t y = top; y < bottom; y++)
        for (int32_t x = left; x < right; x++)
            ow[(int64_t)y * w + x] = tap_sum(tw + x, w, y, top, bottom, kw, 3);
    memset(h_eff, 0, sizeof(float) * (size_t)n);
    for (int32_t y = top; y < bottom; y++)
        for (int32_t x = left; x < right; x++) {
            const int64_t i = (int64_t)y * w + x;
            h_eff[i] = 0.90f * on[i] + 0.10f * ow[i];
        }
    free(lin);
    free(tn);
    free(on);
    free(tw);
    free(ow);
}

/* tile extent along one axis (image_pipeline.rs:400-410) */
static void tile_span(int32_t t, int32_t step, int32_t pad, int32_t size, int32_t *start, int32_t *len)
{
    const int32_t x0 = t * step;
    const int32_t s = x0 > pad ? x0 - pad : 0;
    const int32_t e = x0 + step + pad < size ? x0 + step + pad : size;
    *start = s;
    *len = e - s;
}

/* Hann weight of tile t at coordinate g (image_pipeline.rs:614-636) */
static float hann(int32_t t, int32_t nt, int32_t step, int32_t pad, int32_t g)
{
    const int32_t x0 = t * step;
    const float pad2 = (float)(2 * pad);
    if (t > 0 && g < x0 + pad) {
        float u = ((float)g - ((float)x0 - (float)pad)) / pad2;
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
        return 0.5f - 0.5f * cosf(PI_F * u);
    }
    if (t + 1 < nt && g >= x0 + step - pad) {
        float u = ((float)g - ((float)(x0 + step) - (float)pad)) / pad2;
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
        return 0.5f + 0.5f * cosf(PI_F * u);
    }
    return 1.0f;
}

static float smooth_up(float v)   /* shadow gate (:759-767, :868-874) */
{
    if (v <= 0.005f) return 0.0f;
    if (v >= 0.14f) return 1.0f;
    const float t = (v - 0.005f) / (0.14f - 0.005f);
    return t * t * (3.0f - 2.0f * t);
}

static float smooth_down(float v) /* highlight gate (:773-780, :877-884) */
{
    if (v <= 0.65f) return 1.0f;
    if (v >= 0.94f) return 0.0f;
    const float u = (v - 0.65f) / (0.94f - 0.65f);
    return 1.0f - u * u * (3.0f - 2.0f * u);
}

int64_t orc_film_tile_values(int32_t w, int32_t h, int32_t step, int32_t pad)
{
    const int32_t nx = (w + step - 1) / step, ny = (h + step - 1) / step;
    int64_t n = 0;
    for (int32_t ty = 0; ty < ny; ty++)
        for (int32_t tx = 0; tx < nx; tx++) {
            int32_t xs, tw, ys, th;
            tile_span(tx, step, pad, w, &xs, &tw);
            tile_span(ty, step, pad, h, &ys, &th);
            n += (int64_t)tw * th;
        }
    return n;
}

/* compositing of the rendered tiles (image_pipeline.rs:593-891).  tiles: D_neg of every
 * tile in tile order (row-major tiles, each tw x th row-major).  Optional outputs
 * (NULL to skip): d_norm, d_low, grain (before the OTF), scanned (after it), tone. */
int orc_film_composite(const orc_film_cfg *c, const uint8_t *src, const float *h_eff, const float *tiles,
                       uint8_t *out, float *d_norm_o, float *d_low_o, float *grain_o, float *scanned_o,
                       float *tone_o, orc_film_stats *st)
{
    const int32_t W = c->width, H = c->height, step = c->tile_step, pad = c->pad;
    if (W <= 0 || H <= 0 || step <= 0 || pad < 0 || 2 * pad > step) return -1;
    const int64_t n = (int64_t)W * H;
    const int32_t nx = (W + step - 1) / step, ny = (H + step - 1) / step;
    float *acc = (float *)calloc((size_t)n, sizeof(float));
    float *wacc = (float *)calloc((size_t)n, sizeof(float));
    float *wsq = (float *)calloc((size_t)n, sizeof(float));
    /* Hann accumulation in tile order (:600-649) */
    const float *tp = tiles;
    for (int32_t ty = 0; ty < ny; ty++)
        for (int32_t tx = 0; tx < nx; tx++) {
            int32_t xs, tw, ys, th;
            tile_span(tx, step, pad, W, &xs, &tw);
            tile_span(ty, step, pad, H, &ys, &th);
            for (int32_t py = 0; py < th; py++) {
                const int32_t gy = ys + py;
                const float wy = hann(ty, ny, step, pad, gy);
                for (int32_t px = 0; px < tw; px++) {
                    const int32_t gx = xs + px;
                    const float wx = hann(tx, nx, step, pad, gx);
                    const float wt = wx * wy;
                    const int64_t gi = (int64_t)gy * W + gx;
                    acc[gi] += tp[(int64_t)py * tw + px] * wt;
                    wacc[gi] += wt;
                    wsq[gi] += wt * wt;
                }
            }
            tp += (int64_t)tw * th;
        }
    float *dn = (float *)malloc(sizeof(float) * (size_t)n);
    for (int64_t i = 0; i < n; i++) dn[i] = acc[i] / fmaxf(wacc[i], 1e-5f);
    /* low-frequency separation (:651-706) */
    const float sigma = fmaxf((float)c->grain_size_px * 1.5f, 1.5f);
    float kr = ceilf(2.5f * sigma);
    kr = kr < 2.0f ? 2.0f : (kr > 48.0f ? 48.0f : kr);
    const int32_t rad = (int32_t)kr;
    float kern[97];
    float ksum = 0.0f;
    for (int32_t i = -rad; i <= rad; i++) {
        const float q = (float)i / sigma;
        const float wv = expf(-0.5f * (q * q));
        kern[i + rad] = wv;
        ksum += wv;
    }
    for (int32_t i = 0; i <= 2 * rad; i++) kern[i] /= fmaxf(ksum, 1e-6f);
    float *tmp = (float *)malloc(sizeof(float) * (size_t)n);
    float *lo = (float *)malloc(sizeof(float) * (size_t)n);
    for (int32_t y = 0; y < H; y++)
        for (int32_t x = 0; x < W; x++) tmp[(int64_t)y * W + x] = tap_sum(dn + (int64_t)y * W, 1, x, 0, W, kern, rad);
    for (int32_t y = 0; y < H; y++)
        for (int32_t x = 0; x < W; x++) lo[(int64_t)y * W + x] = tap_sum(tmp + x, W, y, 0, H, kern, rad);
    /* statistics of the high-pass field (:717-724) */
    const float nf = (float)(n > 0 ? n : 1);
    float mean, sd;
    if (c->fold == 0) {
        float s = 0.0f;
        for (int64_t i = 0; i < n; i++) s += dn[i] - lo[i];
        mean = s / nf;
        float s2 = 0.0f;
        for (int64_t i = 0; i < n; i++) {
            const float d = dn[i] - lo[i] - mean;
            s2 += d * d;
        }
        sd = fmaxf(sqrtf(s2 / nf), 0.035f);
    } else {
        double s = 0.0, s2 = 0.0;
        for (int64_t i = 0; i < n; i++) {
            const float d = dn[i] - lo[i];
            s += (double)d;
            s2 += (double)d * (double)d;
        }
        mean = (float)(s / (double)n);
        const double m = (double)mean;
        const double var = (s2 - 2.0 * m * s + (double)n * m * m) / (double)n;
        sd = fmaxf(sqrtf((float)(var > 0.0 ? var : 0.0)), 0.035f);
    }
    const float push_mult = 1.0f + 0.15f * c->push;
    const float amp = c->base_nominal * push_mult;
    const float norm = fminf((amp * c->strength) / sd, 2.2f);
    /* per-pixel grain and tone (:734-819) */
    float *grain = (float *)malloc(sizeof(float) * (size_t)n);
    float *tone = (float *)malloc(sizeof(float) * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const float wv = fmaxf(wacc[i], 1e-5f);
        const float wq = fmaxf(wsq[i] / (wv * wv), 0.05f);
        const float gscale = fminf(1.0f / sqrtf(wq), 1.45f);
        const float dh = dn[i] - lo[i];
        const float raw = (float)src[i] / 255.0f;
        const float base = raw <= 0.005f ? 0.0f : 0.35f * raw + 0.65f * h_eff[i];
        const float sg = smooth_up(base), hg = smooth_down(base);
        const float bc = base < 0.0f ? 0.0f : (base > 1.0f ? 1.0f : base);
        const float mid = fmaxf(sinf(PI_F * powf(bc, 0.72f)), 0.0f);
        const float vis = sg * hg * mid;
        float tv = base;
        if (c->push > 0.0f) {
            const float gamma = 1.0f + 0.06f * c->push;
            tv = base <= 0.5f ? 0.5f * powf(2.0f * base, gamma) : 1.0f - 0.5f * powf(2.0f * (1.0f - base), gamma);
        }
        const float dd = dh - mean;
        const float ds = fabsf(dd) > 1e-6f ? copysignf(1.0f, dd) * powf(fabsf(dd), 0.85f) : 0.0f;
        grain[i] = ds * norm * gscale * vis;
        tone[i] = tv;
    }
    /* scanner OTF (:822-855) */
    float *scan = grain;
    float *filt = NULL;
    if (c->scanner_otf) {
        filt = (float *)malloc(sizeof(float) * (size_t)n);
        orc_otf3(grain, W, H, filt);
        scan = filt;
    }
    /* gates on the raw input and 8-bit output (:857-891) */
    for (int64_t i = 0; i < n; i++) {
        const float raw = (float)src[i] / 255.0f;
        const float gated = scan[i] * smooth_up(raw) * smooth_down(raw);
        float v = (tone[i] + 0.0f) + gated;
        v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        out[i] = (uint8_t)roundf(v * 255.0f);
    }
    if (st) {
        st->high_mean = mean;
        st->high_std = sd;
        st->norm_factor = norm;
        st->filter_sigma = sigma;
        st->k_rad = rad;
        st->tiles_x = nx;
        st->tiles_y = ny;
    }
    if (d_norm_o) memcpy(d_norm_o, dn, sizeof(float) * (size_t)n);
    if (d_low_o) memcpy(d_low_o, lo, sizeof(float) * (size_t)n);
    if (grain_o) memcpy(grain_o, grain, sizeof(float) * (size_t)n);
    if (scanned_o) memcpy(scanned_o, scan, sizeof(float) * (size_t)n);
    if (tone_o) memcpy(tone_o, tone, sizeof(float) * (size_t)n);
    free(acc);
    free(wacc);
    free(wsq);
    free(dn);
    free(tmp);
    free(lo);
    free(grain);
    free(tone);
    free(filt);
    return 0;
}
