// Full-frame black-and-white compositing kernels (image_pipeline.rs:133-891).
//
// Every float operation follows the reference's f32 order (the library is built with
// -ffp-contract=off and IEEE division and square root), so the optics, the Hann blend
// and the Gaussian low-pass reproduce the reference's buffers bit for bit; the grain
// field differs only by the device's sinf/powf rounding.  The reference's statistics are
// sequential f32 folds over the frame; here they are f64 sums of the same f32 terms,
// reduced per block and then in a fixed order (deterministic).
//
// Memory: every stage streams the frame once (u8 input, f32 buffers), so each kernel is
// HBM-bound; halos come from LDS tiles, not from re-reads:
//   k_film_optics  64 x 16 output tile, 3-pixel halo: 1 B in, 4 B out per pixel
//   k_film_blend   one pixel per lane from up to 4 overlapping tiles
//   k_gauss_h      256-pixel row segments, radius halo
//   k_gauss_v      64 x 64 tiles, radius halo rows in LDS; writes d_high and f64 partial sums
//   k_film_final   64 x 16 tile: grain for the tile and a 1-pixel halo, OTF in LDS, u8 out
#include <hip/hip_runtime.h>

#include "film.h"

namespace fg {

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr int kOX = 64, kOY = 16;     // final output tile
constexpr int kPY = 32;               // optics output tile rows (64 x 32)
constexpr int kVX = 64, kVY = 64;     // vertical Gaussian tile
constexpr int kHX = 256;              // horizontal Gaussian segment
constexpr int kMaxRad = 48;

// ------------------------------------------------------------------ gate and optics
__device__ __forceinline__ int wave_min(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// bounding box of the nonzero pixels: rows and columns outside it are the camera gate's
// all-zero margins (image_pipeline.rs:137-153).  16 pixels per lane (one 16-byte load),
// wave then block min/max, one partial box per block (part[4 b .. 4 b + 3]); no atomics.
__global__ void __launch_bounds__(256) k_bbox(const uint8_t *gray, uint32_t W, uint32_t n, int32_t *part)
{
    __shared__ int red[4][4];
    const uint32_t i0 = (blockIdx.x * 256u + threadIdx.x) * 16u;
    int rmin = INT_MAX, rmax = -1, cmin = INT_MAX, cmax = -1;
    if (i0 < n) {
        uint8_t v[16];
        if (i0 + 16u <= n) {
            const uint4 q = *reinterpret_cast<const uint4 *>(gray + i0);
            const uint32_t w[4] = {q.x, q.y, q.z, q.w};
            for (int k = 0; k < 16; k++) v[k] = (uint8_t)(w[k >> 2] >> (8 * (k & 3)));
        } else {
            for (int k = 0; k < 16; k++) v[k] = i0 + k < n ? gray[i0 + k] : 0;
        }
        uint32_t r = i0 / W, c = i0 - r * W;
        for (int k = 0; k < 16; k++) {
            if (v[k]) {
                rmin = min(rmin, (int)r);
                rmax = max(rmax, (int)r);
                cmin = min(cmin, (int)c);
                cmax = max(cmax, (int)c);
            }
            if (++c == W) {
                c = 0;
                r++;
            }
        }
    }
    rmin = wave_min(rmin);
    rmax = wave_max(rmax);
    cmin = wave_min(cmin);
    cmax = wave_max(cmax);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = rmin;
        red[1][wv] = rmax;
        red[2][wv] = cmin;
        red[3][wv] = cmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t *o = part + 4u * blockIdx.x;
        o[0] = min(min(red[0][0], red[0][1]), min(red[0][2], red[0][3]));
        o[1] = max(max(red[1][0], red[1][1]), max(red[1][2], red[1][3]));
        o[2] = min(min(red[2][0], red[2][1]), min(red[2][2], red[2][3]));
        o[3] = max(max(red[3][0], red[3][1]), max(red[3][2], red[3][3]));
    }
}

// the frame's box from the per-block boxes: bbox = (first row, last row, first column,
// last column) of the nonzero pixels, (H, -1, W, -1) when there are none
__global__ void __launch_bounds__(1024) k_bbox_reduce(const int32_t *part, uint32_t nb, uint32_t W, uint32_t H,
                                                      int32_t *bbox)
{
    __shared__ int red[4][16];
    int rmin = INT_MAX, rmax = -1, cmin = INT_MAX, cmax = -1;
    for (uint32_t b = threadIdx.x; b < nb; b += 1024u) {
        rmin = min(rmin, part[4 * b]);
        rmax = max(rmax, part[4 * b + 1]);
        cmin = min(cmin, part[4 * b + 2]);
        cmax = max(cmax, part[4 * b + 3]);
    }
    rmin = wave_min(rmin);
    rmax = wave_max(rmax);
    cmin = wave_min(cmin);
    cmax = wave_max(cmax);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = rmin;
        red[1][wv] = rmax;
        red[2][wv] = cmin;
        red[3][wv] = cmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 16; k++) {
            red[0][0] = min(red[0][0], red[0][k]);
            red[1][0] = max(red[1][0], red[1][k]);
            red[2][0] = min(red[2][0], red[2][k]);
            red[3][0] = max(red[3][0], red[3][k]);
        }
        const bool any = red[1][0] >= 0;
        bbox[0] = any ? red[0][0] : (int32_t)H;
        bbox[1] = red[1][0];
        bbox[2] = any ? red[2][0] : (int32_t)W;
        bbox[3] = red[3][0];
    }
}

// The filter passes of one tile; kInterior: the tile and its 3-pixel halo lie inside the
// gate, so every tap is inside and the gate tests drop out (same operations, same order).
template <bool kInterior>
__device__ __forceinline__ void optics_tile(float (*hl)[kOX + 6], float (*tn)[kOX], float (*tw)[kOX],
                                            int x0, int y0, int top, int bottom, int left, int right,
                                            uint32_t W, uint32_t H, float *h_eff)
{
    constexpr float KN[3] = {0.15f, 0.70f, 0.15f};
    constexpr float KW[7] = {0.03f, 0.12f, 0.20f, 0.30f, 0.20f, 0.12f, 0.03f};
    const int t = threadIdx.x;
    for (int i = t; i < (kPY + 2) * kOX; i += 256) {
        const int r = i / kOX, c = i % kOX;
        const int y = y0 - 1 + r, x = x0 + c;
        float v = 0.0f;
        if (kInterior || (y >= top && y < bottom && x >= left && x < right)) {
            float sum = 0.0f, ws = 0.0f;
            for (int j = 0; j < 3; j++) {
                const int tap = x + j - 1;
                if (kInterior || (tap >= left && tap < right)) {
                    sum += hl[r + 2][c + 2 + j] * KN[j];
                    ws += KN[j];
                }
            }
            v = sum / fmaxf(ws, 1e-6f);
        }
        tn[r][c] = v;
    }
    for (int i = t; i < (kPY + 6) * kOX; i += 256) {
        const int r = i / kOX, c = i % kOX;
        const int y = y0 - 3 + r, x = x0 + c;
        float v = 0.0f;
        if (kInterior || (y >= top && y < bottom && x >= left && x < right)) {
            float sum = 0.0f, ws = 0.0f;
            for (int j = 0; j < 7; j++) {
                const int tap = x + j - 3;
                if (kInterior || (tap >= left && tap < right)) {
                    sum += hl[r][c + j] * KW[j];
                    ws += KW[j];
                }
            }
            v = sum / fmaxf(ws, 1e-6f);
        }
        tw[r][c] = v;
    }
    __syncthreads();
    for (int i = t; i < kPY * kOX; i += 256) {
        const int r = i / kOX, c = i % kOX;
        const int y = y0 + r, x = x0 + c;
        if (y >= (int)H || x >= (int)W) continue;
        float h = 0.0f;
        if (kInterior || (y >= top && y < bottom && x >= left && x < right)) {
            float sn = 0.0f, wn = 0.0f;
            for (int j = 0; j < 3; j++) {
                const int tap = y + j - 1;
                if (kInterior || (tap >= top && tap < bottom)) {
                    sn += tn[r + j][c] * KN[j];
                    wn += KN[j];
                }
            }
            float sw = 0.0f, ww = 0.0f;
            for (int j = 0; j < 7; j++) {
                const int tap = y + j - 3;
                if (kInterior || (tap >= top && tap < bottom)) {
                    sw += tw[r + j][c] * KW[j];
                    ww += KW[j];
                }
            }
            const float on = sn / fmaxf(wn, 1e-6f), ow = sw / fmaxf(ww, 1e-6f);
            h = 0.90f * on + 0.10f * ow;
        }
        h_eff[(size_t)y * W + x] = h;
    }
}

// apply_camera_and_gelatin_optics (image_pipeline.rs:155-249): narrow [0.15, 0.70, 0.15]
// and wide 7-tap separable passes with taps restricted to the gate and renormalised,
// h = 0.90 narrow + 0.10 wide inside the gate, 0 outside.
__global__ void __launch_bounds__(256) k_film_optics(const uint8_t *gray, uint32_t W, uint32_t H,
                                                     const int32_t *bbox, float *h_eff)
{
    __shared__ float hl[kPY + 6][kOX + 6];
    __shared__ float tn[kPY + 2][kOX];
    __shared__ float tw[kPY + 6][kOX];
    __shared__ float lut[256];   // v / 255 for every 8-bit value, correctly rounded
    const int top = bbox[0], bottom = bbox[1] + 1, left = bbox[2], right = bbox[3] + 1;
    const int x0 = blockIdx.x * kOX, y0 = blockIdx.y * kPY;
    const int t = threadIdx.x;
    lut[t] = (float)t / 255.0f;
    // all of a lane's tile loads are issued before the first is used
    constexpr int kTile = (kPY + 6) * (kOX + 6), kPer = (kTile + 255) / 256;
    uint32_t g8[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int i = t + 256 * j;
        const int r = i / (kOX + 6), c = i % (kOX + 6);
        const int y = y0 - 3 + r, x = x0 - 3 + c;
        g8[j] = 0xFFFFFFFFu;   // outside the gate
        if (i < kTile && y >= top && y < bottom && x >= left && x < right) g8[j] = gray[(size_t)y * W + x];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int i = t + 256 * j;
        if (i < kTile) hl[i / (kOX + 6)][i % (kOX + 6)] = g8[j] == 0xFFFFFFFFu ? 0.0f : lut[g8[j]];
    }
    __syncthreads();
    if (x0 - 3 >= left && x0 + kOX + 3 <= right && y0 - 3 >= top && y0 + kPY + 3 <= bottom)
        optics_tile<true>(hl, tn, tw, x0, y0, top, bottom, left, right, W, H, h_eff);
    else
        optics_tile<false>(hl, tn, tw, x0, y0, top, bottom, left, right, W, H, h_eff);
}

__device__ __forceinline__ uint32_t luma_of(uint32_t r8, uint32_t g8, uint32_t b8, float rw, float gw, float bw)
{
    float l = (float)r8 * rw + (float)g8 * gw + (float)b8 * bw;
    l = fminf(fmaxf(l, 0.0f), 255.0f);
    return (uint32_t)roundf(l);
}

// convert_to_panchromatic_luma (image_pipeline.rs:284-293), 4 pixels per lane: three
// 4-byte loads of interleaved RGB, one 4-byte store; the last n % 4 pixels one by one
__global__ void __launch_bounds__(256) k_film_luma(const uint8_t *rgb, uint64_t n, float rw, float gw, float bw,
                                                   uint8_t *gray)
{
    const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint64_t i = 4 * q;
    if (i + 4 <= n) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(rgb + 3 * i);
        const uint32_t w0 = src[0], w1 = src[1], w2 = src[2];
        const uint32_t b[12] = {w0 & 255u, (w0 >> 8) & 255u, (w0 >> 16) & 255u, w0 >> 24,
                                w1 & 255u, (w1 >> 8) & 255u, (w1 >> 16) & 255u, w1 >> 24,
                                w2 & 255u, (w2 >> 8) & 255u, (w2 >> 16) & 255u, w2 >> 24};
        uint32_t o = 0;
        for (int k = 0; k < 4; k++) o |= luma_of(b[3 * k], b[3 * k + 1], b[3 * k + 2], rw, gw, bw) << (8 * k);
        *reinterpret_cast<uint32_t *>(gray + i) = o;
    } else {
        for (uint64_t p = i; p < n; p++)
            gray[p] = (uint8_t)luma_of(rgb[3 * p], rgb[3 * p + 1], rgb[3 * p + 2], rw, gw, bw);
    }
}

// ------------------------------------------------------------------ Hann blend
// tiles covering coordinate g along one axis, in tile order, with their Hann weights
// (image_pipeline.rs:614-636): a coordinate lies in its home tile and, within the
// overlap, in one neighbour.
__device__ __forceinline__ int axis_cover(uint32_t g, uint32_t nt, uint32_t step, uint32_t pad, const float *ramp,
                                          uint32_t t[2], float w[2])
{
    const uint32_t h = min(g / step, nt - 1u);
    const uint32_t x0 = h * step;
    if (pad) {
        if (h > 0u && g < x0 + pad) {
            const uint32_t k = g + pad - x0;
            t[0] = h - 1u;
            w[0] = ramp[2u * pad + k];
            t[1] = h;
            w[1] = ramp[k];
            return 2;
        }
        if (h + 1u < nt && g >= x0 + step - pad) {
            const uint32_t k = g - (x0 + step - pad);
            t[0] = h;
            w[0] = ramp[2u * pad + k];
            t[1] = h + 1u;
            w[1] = ramp[k];
            return 2;
        }
    }
    t[0] = h;
    w[0] = 1.0f;
    return 1;
}

// d_accum, weight_accum and weight_sq_accum of one pixel (image_pipeline.rs:638-646),
// accumulated in tile order
template <bool kRead>
__device__ __forceinline__ void hann_pixel(const FilmLayout &L, const float *tiles, uint32_t x, uint32_t y,
                                           float &acc, float &wa, float &wq)
{
    uint32_t tx[2], ty[2];
    float wx[2], wy[2];
    const int nx = axis_cover(x, L.ntx, L.step, L.pad, L.ramp, tx, wx);
    const int ny = axis_cover(y, L.nty, L.step, L.pad, L.ramp, ty, wy);
    acc = 0.0f;
    wa = 0.0f;
    wq = 0.0f;
    for (int a = 0; a < ny; a++)
        for (int b = 0; b < nx; b++) {
            const float wt = wx[b] * wy[a];
            if (kRead) {
                const uint32_t Ty = ty[a], Tx = tx[b];
                const uint64_t base = (uint64_t)L.yoff[Ty] * L.wext + (uint64_t)L.th[Ty] * L.xoff[Tx];
                const uint64_t local = (uint64_t)(y - L.ys[Ty]) * L.tw[Tx] + (x - L.xs[Tx]);
                acc += tiles[base + local] * wt;
            }
            wa += wt;
            wq += wt * wt;
        }
}

__global__ void __launch_bounds__(256) k_film_blend(FilmLayout L, const float *tiles, float *d_norm)
{
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= L.W) return;
    float acc, wa, wq;
    hann_pixel<true>(L, tiles, x, y, acc, wa, wq);
    d_norm[(size_t)y * L.W + x] = acc / fmaxf(wa, 1e-5f);
}

// ------------------------------------------------------------------ Gaussian low-pass
// horizontal pass over one row segment (image_pipeline.rs:672-687)
__global__ void __launch_bounds__(256) k_gauss_h(const float *src, uint32_t W, const float *kern, int rad,
                                                 float *dst)
{
    __shared__ float seg[kHX + 2 * kMaxRad];
    __shared__ float kk[2 * kMaxRad + 1];
    const int y = blockIdx.y, x0 = blockIdx.x * kHX, t = threadIdx.x;
    const float *row = src + (size_t)y * W;
    for (int i = t; i < kHX + 2 * rad; i += 256) {
        const int x = x0 - rad + i;
        seg[i] = (x >= 0 && x < (int)W) ? row[x] : 0.0f;
    }
    for (int i = t; i <= 2 * rad; i += 256) kk[i] = kern[i];
    __syncthreads();
    const int x = x0 + t;
    if (x >= (int)W) return;
    float sum = 0.0f, ws = 0.0f;
    for (int j = 0; j <= 2 * rad; j++) {
        const int tap = x + j - rad;
        if (tap >= 0 && tap < (int)W) {
            sum += seg[t + j] * kk[j];
            ws += kk[j];
        }
    }
    dst[(size_t)y * W + x] = sum / fmaxf(ws, 1e-6f);
}

// vertical pass (image_pipeline.rs:689-704), then d_high = d_norm - d_low and per-block
// f64 sums of d_high and d_high^2 (partials[2 b], partials[2 b + 1])
__global__ void __launch_bounds__(256) k_gauss_v(const float *tmp, const float *d_norm, uint32_t W, uint32_t H,
                                                 const float *kern, int rad, float *d_high, double *partials)
{
    extern __shared__ float col[];   // (kVY + 2 rad) rows x kVX columns
    __shared__ float kk[2 * kMaxRad + 1];
    __shared__ double red[2][4];
    const int x0 = blockIdx.x * kVX, y0 = blockIdx.y * kVY, t = threadIdx.x;
    const int rows = kVY + 2 * rad;
    for (int i = t; i < rows * kVX; i += 256) {
        const int r = i / kVX, c = i % kVX;
        const int y = y0 - rad + r, x = x0 + c;
        col[i] = (y >= 0 && y < (int)H && x < (int)W) ? tmp[(size_t)y * W + x] : 0.0f;
    }
    for (int i = t; i <= 2 * rad; i += 256) kk[i] = kern[i];
    __syncthreads();
    const int c = t % kVX, x = x0 + c;
    double s1 = 0.0, s2 = 0.0;
    for (int r = t / kVX; r < kVY; r += 256 / kVX) {
        const int y = y0 + r;
        if (y >= (int)H || x >= (int)W) continue;
        float sum = 0.0f, ws = 0.0f;
        for (int j = 0; j <= 2 * rad; j++) {
            const int tap = y + j - rad;
            if (tap >= 0 && tap < (int)H) {
                sum += col[(r + j) * kVX + c] * kk[j];
                ws += kk[j];
            }
        }
        const float lo = sum / fmaxf(ws, 1e-6f);
        const size_t idx = (size_t)y * W + x;
        const float dh = d_norm[idx] - lo;
        d_high[idx] = dh;
        s1 += (double)dh;
        s2 += (double)dh * (double)dh;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if ((t & 63) == 0) {
        red[0][t >> 6] = s1;
        red[1][t >> 6] = s2;
    }
    __syncthreads();
    if (t == 0) {
        const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        partials[2 * b] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        partials[2 * b + 1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    }
}

// high_mean, high_std and the grain normalisation (image_pipeline.rs:717-731) from the
// per-block sums, reduced in a fixed order
__global__ void __launch_bounds__(1024) k_film_stats(const double *partials, uint32_t nb, uint64_t n_px,
                                                     float strength, float base_amp, float *fstats)
{
    __shared__ double r1[1024], r2[1024];
    const int t = threadIdx.x;
    double s1 = 0.0, s2 = 0.0;
    for (uint32_t b = t; b < nb; b += 1024) {
        s1 += partials[2 * b];
        s2 += partials[2 * b + 1];
    }
    r1[t] = s1;
    r2[t] = s2;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) {
            r1[t] += r1[t + o];
            r2[t] += r2[t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double n = (double)n_px;
        const float mean = (float)(r1[0] / n);
        const double m = (double)mean;
        const double var = (r2[0] - 2.0 * m * r1[0] + n * m * m) / n;
        const float sd = fmaxf(sqrtf((float)(var > 0.0 ? var : 0.0)), 0.035f);
        fstats[0] = mean;
        fstats[1] = sd;
        fstats[2] = fminf((base_amp * strength) / sd, 2.2f);
        fstats[3] = 0.0f;
    }
}

// ------------------------------------------------------------------ grain and output
__device__ __forceinline__ float shadow_gate(float v)   // image_pipeline.rs:759-767, :868-874
{
    if (v <= 0.005f) return 0.0f;
    if (v >= 0.14f) return 1.0f;
    const float t = (v - 0.005f) / (0.14f - 0.005f);
    return t * t * (3.0f - 2.0f * t);
}

__device__ __forceinline__ float highlight_gate(float v)   // :773-780, :877-884
{
    if (v <= 0.65f) return 1.0f;
    if (v >= 0.94f) return 0.0f;
    const float u = (v - 0.65f) / (0.94f - 0.65f);
    return 1.0f - u * u * (3.0f - 2.0f * u);
}

__device__ __forceinline__ float base_value(uint8_t g8, float h)   // :743-753
{
    const float raw = (float)g8 / 255.0f;
    return raw <= 0.005f ? 0.0f : 0.35f * raw + 0.65f * h;
}

// grain_buf of one pixel (image_pipeline.rs:736-817)
__device__ __forceinline__ float grain_value(const FilmLayout &L, const uint8_t *gray, const float *h_eff,
                                             const float *d_high, const float *fs, uint32_t x, uint32_t y)
{
    float acc, wa, wq;
    hann_pixel<false>(L, nullptr, x, y, acc, wa, wq);
    const float wv = fmaxf(wa, 1e-5f);
    const float wn = fmaxf(wq / (wv * wv), 0.05f);
    const float gscale = fminf(1.0f / sqrtf(wn), 1.45f);
    const size_t idx = (size_t)y * L.W + x;
    const float base = base_value(gray[idx], h_eff[idx]);
    const float bc = fminf(fmaxf(base, 0.0f), 1.0f);
    const float mid = fmaxf(sinf(kPi * powf(bc, 0.72f)), 0.0f);
    const float vis = shadow_gate(base) * highlight_gate(base) * mid;
    const float dd = d_high[idx] - fs[0];
    const float ds = fabsf(dd) > 1e-6f ? copysignf(1.0f, dd) * powf(fabsf(dd), 0.85f) : 0.0f;
    return ds * fs[2] * gscale * vis;
}

// grain for the tile and a clamp-to-edge halo, scanner OTF [0.035, 0.93, 0.035] in LDS
// (image_pipeline.rs:822-855), gates on the raw input, tone curve and 8-bit output
// (:786-801, :857-891)
template <bool kOtf>
__global__ void __launch_bounds__(256) k_film_final(FilmLayout L, FilmParams P, const uint8_t *gray,
                                                    const float *h_eff, const float *d_high, const float *fstats,
                                                    float *grain_out, uint8_t *out)
{
    __shared__ float g[kOY + 2][kOX + 2];
    __shared__ float th[kOY + 2][kOX];
    __shared__ float fs[4];
    const int x0 = blockIdx.x * kOX, y0 = blockIdx.y * kOY, t = threadIdx.x;
    if (t < 4) fs[t] = fstats[t];
    __syncthreads();
    const int W = (int)L.W, H = (int)L.H;
    for (int i = t; i < (kOY + 2) * (kOX + 2); i += 256) {
        const int r = i / (kOX + 2), c = i % (kOX + 2);
        if (!kOtf && (r == 0 || c == 0 || r == kOY + 1 || c == kOX + 1)) continue;
        const int y = min(max(y0 - 1 + r, 0), H - 1), x = min(max(x0 - 1 + c, 0), W - 1);
        g[r][c] = grain_value(L, gray, h_eff, d_high, fs, (uint32_t)x, (uint32_t)y);
    }
    __syncthreads();
    constexpr float k0 = 0.035f, k1 = 0.93f, k2 = 0.035f;
    if (kOtf) {
        for (int i = t; i < (kOY + 2) * kOX; i += 256) {
            const int r = i / kOX, c = i % kOX;
            th[r][c] = g[r][c] * k0 + g[r][c + 1] * k1 + g[r][c + 2] * k2;
        }
        __syncthreads();
    }
    for (int i = t; i < kOY * kOX; i += 256) {
        const int r = i / kOX, c = i % kOX;
        const int y = y0 + r, x = x0 + c;
        if (y >= H || x >= W) continue;
        const float sc = kOtf ? th[r][c] * k0 + th[r + 1][c] * k1 + th[r + 2][c] * k2 : g[r + 1][c + 1];
        const size_t idx = (size_t)y * W + x;
        const uint8_t g8 = gray[idx];
        const float raw = (float)g8 / 255.0f;
        const float base = raw <= 0.005f ? 0.0f : 0.35f * raw + 0.65f * h_eff[idx];
        float tone = base;
        if (P.push > 0.0f) {
            const float gamma = 1.0f + 0.06f * P.push;
            tone = base <= 0.5f ? 0.5f * powf(2.0f * base, gamma) : 1.0f - 0.5f * powf(2.0f * (1.0f - base), gamma);
        }
        const float gated = sc * shadow_gate(raw) * highlight_gate(raw);
        float v = (tone + 0.0f) + gated;
        v = fminf(fmaxf(v, 0.0f), 1.0f);
        out[idx] = (uint8_t)roundf(v * 255.0f);
        if (grain_out) grain_out[idx] = sc;
    }
}

inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

}  // namespace

size_t film_bbox_size(uint32_t W, uint32_t H) { return 4u + 4u * (size_t)(((uint64_t)W * H + 4095) / 4096); }

hipError_t launch_film_optics(const uint8_t *gray, uint32_t W, uint32_t H, int32_t *bbox, float *h_eff,
                              hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    const uint64_t n = (uint64_t)W * H;
    if (n == 0) return hipSuccess;
    if (n > 0x80000000ull) return hipErrorInvalidValue;
    const uint32_t nb = (uint32_t)((n + 4095) / 4096);
    hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(256), 0, s, gray, W, (uint32_t)n, bbox + 4);
    hipLaunchKernelGGL(k_bbox_reduce, dim3(1), dim3(1024), 0, s, bbox + 4, nb, W, H, bbox);
    hipLaunchKernelGGL(k_film_optics, dim3(cdiv(W, kOX), cdiv(H, kPY)), dim3(256), 0, s, gray, W, H, bbox, h_eff);
    return hipGetLastError();
}

hipError_t launch_film_luma(const uint8_t *rgb, uint64_t n, float rw, float gw, float bw, uint8_t *gray,
                            hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n == 0) return hipSuccess;
    const uint64_t quads = (n + 3) / 4;
    hipLaunchKernelGGL(k_film_luma, dim3((uint32_t)((quads + 255) / 256)), dim3(256), 0, s, rgb, n, rw, gw, bw, gray);
    return hipGetLastError();
}

hipError_t launch_film_blend(const FilmLayout &L, const float *tiles, float *d_norm, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (!L.W || !L.H) return hipSuccess;
    hipLaunchKernelGGL(k_film_blend, dim3(cdiv(L.W, 256), L.H), dim3(256), 0, s, L, tiles, d_norm);
    return hipGetLastError();
}

size_t film_partials_size(uint32_t W, uint32_t H) { return 2ull * cdiv(W, kVX) * cdiv(H, kVY); }

hipError_t launch_film_composite(const FilmLayout &L, const FilmParams &P, const uint8_t *gray,
                                 const float *h_eff, const float *tiles, float *d_norm, float *tmp,
                                 float *d_high, double *partials, float *fstats, float *grain_out,
                                 uint8_t *out, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (!L.W || !L.H) return hipSuccess;
    if (P.rad > (uint32_t)kMaxRad) return hipErrorInvalidValue;
    const int rad = (int)P.rad;
    hipLaunchKernelGGL(k_film_blend, dim3(cdiv(L.W, 256), L.H), dim3(256), 0, s, L, tiles, d_norm);
    hipLaunchKernelGGL(k_gauss_h, dim3(cdiv(L.W, kHX), L.H), dim3(256), 0, s, d_norm, L.W, P.kern, rad, tmp);
    const dim3 vg(cdiv(L.W, kVX), cdiv(L.H, kVY));
    const size_t lds = (size_t)(kVY + 2 * rad) * kVX * sizeof(float);
    hipLaunchKernelGGL(k_gauss_v, vg, dim3(256), lds, s, tmp, d_norm, L.W, L.H, P.kern, rad, d_high, partials);
    hipLaunchKernelGGL(k_film_stats, dim3(1), dim3(1024), 0, s, partials, vg.x * vg.y,
                       (uint64_t)L.W * L.H, P.strength, P.base_amp, fstats);
    const dim3 fg(cdiv(L.W, kOX), cdiv(L.H, kOY));
    if (P.otf)
        hipLaunchKernelGGL(k_film_final<true>, fg, dim3(256), 0, s, L, P, gray, h_eff, d_high, fstats, grain_out,
                           out);
    else
        hipLaunchKernelGGL(k_film_final<false>, fg, dim3(256), 0, s, L, P, gray, h_eff, d_high, fstats, grain_out,
                           out);
    return hipGetLastError();
}

}  // namespace fg
