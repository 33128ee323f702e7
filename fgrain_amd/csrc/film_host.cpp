// C ABI of the full-frame black-and-white film render, render_film_black_and_white_tiled
// (image_pipeline.rs:299-904): panchromatic luma, camera and gelatin optics, the
// overlapping tiles as one batched frame, Hann blend, low-frequency separation,
// sensitometric compositing, scanner OTF and 8-bit output, all on the device.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "engine_state.h"

namespace {

using Clock = std::chrono::steady_clock;

// The compositing kernels (Hann blend, Gaussian separation, statistics, grain and
// output) are switched off: their first launch on the MI355X ended in a device memory
// fault whose cause is not yet found.  The entry points that launch them return
// FG_ERR_PIPELINE until the cause is fixed.  FG_FILM_COMPOSITE=1 at compile time switches
// them on (film.h); the CPU harness in tests/ builds this file and the kernels that way
// for the host, under AddressSanitizer.
constexpr bool kCompositeEnabled = FG_FILM_COMPOSITE != 0;
constexpr const char *kCompositeOff =
    "the compositing kernels are switched off in this build (device memory fault under investigation)";

float ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<float, std::milli>(b - a).count();
}

std::string lower(const char *s)
{
    std::string r(s ? s : "");
    for (auto &c : r) c = (char)tolower((unsigned char)c);
    return r;
}

// spectral weights of convert_to_panchromatic_luma (image_pipeline.rs:264-283)
void luma_weights(const std::string &p, float &rw, float &gw, float &bw)
{
    if (p == "tri-x" || p == "trix" || p == "tri-x-400" || p == "tri-x-1600" || p == "trix-1600" ||
        p == "caffenol") {
        rw = 0.42f; gw = 0.46f; bw = 0.12f;
    } else if (p == "hp5" || p == "hp5-plus") {
        rw = 0.36f; gw = 0.48f; bw = 0.16f;
    } else if (p == "t-max" || p == "tmax") {
        rw = 0.33f; gw = 0.52f; bw = 0.15f;
    } else {
        rw = 0.38f; gw = 0.48f; bw = 0.14f;
    }
}

// base grain amplitude per preset (image_pipeline.rs:726-731)
float base_nominal(const std::string &p)
{
    if (p == "t-max" || p == "tmax") return 0.024f;
    if (p == "hp5" || p == "hp5-plus") return 0.032f;
    if (p == "tri-x-1600" || p == "trix-1600" || p == "caffenol") return 0.042f;
    return 0.038f;
}

bool valid_options(const fg_film_options *o, std::string &msg)
{
    if (!o) {
        msg = "null options";
        return false;
    }
    if (memchr(o->preset, 0, sizeof o->preset) == nullptr) {
        msg = "preset is not NUL-terminated";
        return false;
    }
    if (!(o->grain_size_px > 0) || !std::isfinite(o->grain_size_px) || !std::isfinite(o->grain_strength) ||
        !std::isfinite(o->exposure) || !std::isfinite(o->scale) || !std::isfinite(o->push)) {
        msg = "grain size must be positive and all values finite";
        return false;
    }
    if (o->tile_size < 48) {  // u32::clamp(48, tile_size) panics in the reference
        msg = "tile_size must be at least 48";
        return false;
    }
    if (o->synthesis > 1 || (o->frame_mode != FG_FRAME_PLAIN && o->frame_mode != FG_FRAME_BATCHED)) {
        msg = "synthesis must be 0 or 1 and frame_mode FG_FRAME_PLAIN or FG_FRAME_BATCHED";
        return false;
    }
    return true;
}

// image_pipeline.rs:346-385 and :651-658
void make_plan(const fg_film_options &o, uint32_t W, uint32_t H, fg_film_plan &p)
{
    memset(&p, 0, sizeof p);
    const std::string pr = lower(o.preset);
    const double user_scale = o.scale > 0.0 ? o.scale : 1.0;
    double res_scale = 1.0, selwyn = 1.0;
    if (o.auto_scale) {
        const double ref = std::max((double)o.reference_res, 512.0);
        const double long_edge = std::max((double)std::max(W, H), 1.0);
        const double ratio = std::min(std::max(long_edge / ref, 0.05), 5.0);
        res_scale = ratio;
        selwyn = sqrt(ratio);
    }
    p.grain_size_px = std::min(std::max(o.grain_size_px * (user_scale * res_scale), 0.05), 15.0);
    p.strength = std::max(o.grain_strength * sqrt(user_scale) * selwyn, 0.0);
    p.pixel_size_um = 1.0 / p.grain_size_px;
    p.push = (pr.find("1600") != std::string::npos || pr == "caffenol") ? std::max(o.push, 2.0)
                                                                       : std::max(o.push, 0.0);
    const double st = std::min(std::max(round(384.0 * p.grain_size_px), 0.0), 4294967295.0);
    const uint32_t step = std::max(std::min(std::max((uint32_t)st, 48u), o.tile_size), 48u);
    p.tile_step = step;
    p.pad = std::min(std::max(step / 10u, 12u), 48u);
    p.tiles_x = (W + step - 1) / step;
    p.tiles_y = (H + step - 1) / step;
    p.filter_sigma = fmaxf((float)p.grain_size_px * 1.5f, 1.5f);
    p.k_rad = (uint32_t)fminf(fmaxf(ceilf(2.5f * p.filter_sigma), 2.0f), 48.0f);
    p.base_nominal = base_nominal(pr);
}

// normalised Gaussian of the low-frequency separation (image_pipeline.rs:659-668)
std::vector<float> gauss_kernel(float sigma, int rad)
{
    std::vector<float> k;
    float sum = 0.0f;
    for (int i = -rad; i <= rad; i++) {
        const float q = (float)i / sigma;
        const float w = expf(-0.5f * (q * q));
        k.push_back(w);
        sum += w;
    }
    for (auto &w : k) w /= fmaxf(sum, 1e-6f);
    return k;
}

int64_t tile_values(uint32_t W, uint32_t H, uint32_t step, uint32_t pad)
{
    int64_t n = 0;
    for (uint32_t ty = 0; ty < (H + step - 1) / step; ty++)
        for (uint32_t tx = 0; tx < (W + step - 1) / step; tx++) {
            uint32_t xs, tw, ys, th;
            fg::film_span(tx, step, pad, W, xs, tw);
            fg::film_span(ty, step, pad, H, ys, th);
            n += (int64_t)tw * th;
        }
    return n;
}

bool frame_ok(uint32_t W, uint32_t H, std::string &msg)
{
    if (W == 0 || H == 0) {
        msg = "empty image";
        return false;
    }
    if ((uint64_t)W * H > (1ull << 31)) {
        msg = "image above 2^31 pixels";
        return false;
    }
    return true;
}

// compositing on device buffers; fstats left in E->film_fstats
fg_status composite_device(fg_engine *E, const fg_film_plan &plan, uint32_t otf, const fg::FilmLayout &L,
                           const uint8_t *d_gray, const float *d_heff, const float *d_tiles, bool want_grain,
                           float *ms_kernels)
{
    const size_t n = (size_t)L.W * L.H;
    const std::vector<float> k = gauss_kernel(plan.filter_sigma, (int)plan.k_rad);
    HIP_TRY(E->film_kern.ensure(k.size() * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_dnorm.ensure(n * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_tmp.ensure(n * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_dhigh.ensure(n * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_partials.ensure(fg::film_partials_size(L.W, L.H) * 8), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_fstats.ensure(16), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_out.ensure(n), FG_ERR_BUFFER_RANGE);
    if (want_grain) HIP_TRY(E->film_grain.ensure(n * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemcpyAsync(E->film_kern.p, k.data(), k.size() * 4, hipMemcpyHostToDevice, E->stream),
            FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);   // start event on an idle stream
    fg::FilmParams P{};
    P.strength = (float)plan.strength;
    P.base_amp = plan.base_nominal * (1.0f + 0.15f * (float)plan.push);
    P.push = (float)plan.push;
    P.otf = otf;
    P.rad = plan.k_rad;
    P.kern = E->film_kern.as<float>();
    HIP_TRY(hipEventRecord(E->ev0, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(fg::launch_film_composite(L, P, d_gray, d_heff, d_tiles, E->film_dnorm.as<float>(),
                                      E->film_tmp.as<float>(), E->film_dhigh.as<float>(),
                                      E->film_partials.as<double>(), E->film_fstats.as<float>(),
                                      want_grain ? E->film_grain.as<float>() : nullptr,
                                      E->film_out.as<uint8_t>(), E->stream),
            FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev1, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(hipEventSynchronize(E->ev1), FG_ERR_CHANNEL);
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, E->ev0, E->ev1), FG_ERR_CHANNEL);
    if (ms_kernels) *ms_kernels = ms;
    return FG_OK;
}

// gray image on the device: upload (and luma for RGB) into E->film_gray
fg_status upload_gray(fg_engine *E, const uint8_t *pixels, uint32_t W, uint32_t H, uint32_t channels,
                      const std::string &preset)
{
    const size_t n = (size_t)W * H;
    HIP_TRY(E->film_gray.ensure(n), FG_ERR_BUFFER_RANGE);
    if (channels == 1) {
        HIP_TRY(hipMemcpyAsync(E->film_gray.p, pixels, n, hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
        return FG_OK;
    }
    float rw, gw, bw;
    luma_weights(preset, rw, gw, bw);
    HIP_TRY(E->film_rgb.ensure(3 * n), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemcpyAsync(E->film_rgb.p, pixels, 3 * n, hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(fg::launch_film_luma(E->film_rgb.as<uint8_t>(), n, rw, gw, bw, E->film_gray.as<uint8_t>(), E->stream),
            FG_ERR_PIPELINE);
    return FG_OK;
}

}  // namespace

namespace fg {

hipError_t film_layout_upload(uint32_t W, uint32_t H, uint32_t step, uint32_t pad, DevBuf &buf, FilmLayout &L,
                              hipStream_t stream)
{
    const uint32_t ntx = (W + step - 1) / step, nty = (H + step - 1) / step;
    std::vector<uint32_t> u(3 * (size_t)ntx + 3 * (size_t)nty);
    uint32_t *xs = u.data(), *tw = xs + ntx, *xoff = tw + ntx;
    uint32_t *ys = xoff + ntx, *th = ys + nty, *yoff = th + nty;
    uint64_t wext = 0, hext = 0;
    for (uint32_t t = 0; t < ntx; t++) {
        film_span(t, step, pad, W, xs[t], tw[t]);
        xoff[t] = (uint32_t)wext;
        wext += tw[t];
    }
    for (uint32_t t = 0; t < nty; t++) {
        film_span(t, step, pad, H, ys[t], th[t]);
        yoff[t] = (uint32_t)hext;
        hext += th[t];
    }
    if (wext > 0xFFFFFFFFull || hext > 0xFFFFFFFFull) return hipErrorInvalidValue;
    // Hann ramps with the reference's f32 operations (image_pipeline.rs:617-632):
    // t = k / (2 pad) is exact in f32, w = 0.5 -+ 0.5 cos(pi t)
    std::vector<float> ramp(4 * (size_t)pad);
    const float pad2 = (float)(2 * pad);
    for (uint32_t k = 0; k < 2 * pad; k++) {
        float t = (float)k / pad2;
        t = std::min(std::max(t, 0.0f), 1.0f);
        const float c = cosf(3.14159265358979323846f * t);
        ramp[k] = 0.5f - 0.5f * c;
        ramp[2 * pad + k] = 0.5f + 0.5f * c;
    }
    const size_t ub = u.size() * 4, rb = ramp.size() * 4;
    const size_t roff = (ub + 255) & ~(size_t)255;
    hipError_t e = buf.ensure(roff + rb + 4);
    if (e != hipSuccess) return e;
    // on the stream of the kernels that read the tables, complete before the host
    // vectors go out of scope
    e = hipMemcpyAsync(buf.p, u.data(), ub, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && rb)
        e = hipMemcpyAsync(static_cast<char *>(buf.p) + roff, ramp.data(), rb, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    const uint32_t *d = buf.as<uint32_t>();
    L.W = W;
    L.H = H;
    L.step = step;
    L.pad = pad;
    L.ntx = ntx;
    L.nty = nty;
    L.wext = wext;
    L.xs = d;
    L.tw = d + ntx;
    L.xoff = d + 2 * ntx;
    L.ys = d + 3 * ntx;
    L.th = d + 3 * ntx + nty;
    L.yoff = d + 3 * ntx + 2 * nty;
    L.ramp = reinterpret_cast<const float *>(static_cast<const char *>(buf.p) + roff);
    return hipSuccess;
}

}  // namespace fg

extern "C" fg_status fg_film_options_default(fg_film_options *o)
{
    if (!o) return fail(FG_ERR_INVALID_ARG, "fg_film_options_default: null argument");
    memset(o, 0, sizeof *o);
    strcpy(o->preset, "tri-x");
    o->grain_size_px = 2.2;
    o->grain_strength = 1.0;
    o->exposure = 1.0;
    o->scale = 1.0;
    o->push = 0.0;
    o->seed = 42;
    o->tile_size = 320;
    o->reference_res = 6000;
    o->auto_scale = 0;
    o->scanner_otf = 1;
    o->synthesis = 0;
    o->frame_mode = FG_FRAME_BATCHED;
    return FG_OK;
}

extern "C" fg_status fg_film_plan_for(const fg_film_options *o, uint32_t W, uint32_t H, fg_film_plan *plan)
{
    std::string msg;
    if (!plan) return fail(FG_ERR_INVALID_ARG, "fg_film_plan_for: null argument");
    if (!valid_options(o, msg) || !frame_ok(W, H, msg)) return fail(FG_ERR_INVALID_ARG, "fg_film_plan_for: " + msg);
    make_plan(*o, W, H, *plan);
    return FG_OK;
}

extern "C" fg_status fg_film_luma(fg_engine *E, const uint8_t *rgb, uint32_t W, uint32_t H, const char *preset,
                                  uint8_t *gray, float *ms_kernel)
{
    std::string msg;
    if (!E || !rgb || !gray || !preset) return fail(FG_ERR_INVALID_ARG, "fg_film_luma: null argument");
    if (!frame_ok(W, H, msg)) return fail(FG_ERR_INVALID_ARG, "fg_film_luma: " + msg);
    std::lock_guard<std::mutex> fl(E->film_mu);
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    const size_t n = (size_t)W * H;
    float rw, gw, bw;
    luma_weights(lower(preset), rw, gw, bw);
    HIP_TRY(E->film_gray.ensure(n), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_rgb.ensure(3 * n), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemcpyAsync(E->film_rgb.p, rgb, 3 * n, hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    // the kernel's start event on an idle stream: after a staged (pageable) upload the
    // next event would not carry a device timestamp of the kernel's start
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    HIP_TRY(hipEventRecord(E->ev0, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(fg::launch_film_luma(E->film_rgb.as<uint8_t>(), n, rw, gw, bw, E->film_gray.as<uint8_t>(), E->stream),
            FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev1, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(hipMemcpyAsync(gray, E->film_gray.p, n, hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    if (ms_kernel) HIP_TRY(hipEventElapsedTime(ms_kernel, E->ev0, E->ev1), FG_ERR_CHANNEL);
    return FG_OK;
}

extern "C" fg_status fg_film_optics(fg_engine *E, const uint8_t *gray, uint32_t W, uint32_t H, float *h_eff,
                                    float *ms_kernel)
{
    std::string msg;
    if (!E || !gray || !h_eff) return fail(FG_ERR_INVALID_ARG, "fg_film_optics: null argument");
    if (!frame_ok(W, H, msg)) return fail(FG_ERR_INVALID_ARG, "fg_film_optics: " + msg);
    std::lock_guard<std::mutex> fl(E->film_mu);
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    const size_t n = (size_t)W * H;
    fg_status s = upload_gray(E, gray, W, H, 1, "");
    if (s != FG_OK) return s;
    HIP_TRY(E->film_heff.ensure(n * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_bbox.ensure(fg::film_bbox_size(W, H) * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);   // see fg_film_luma
    HIP_TRY(hipEventRecord(E->ev0, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(fg::launch_film_optics(E->film_gray.as<uint8_t>(), W, H, E->film_bbox.as<int32_t>(),
                                   E->film_heff.as<float>(), E->stream),
            FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev1, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(hipMemcpyAsync(h_eff, E->film_heff.p, n * 4, hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    if (ms_kernel) HIP_TRY(hipEventElapsedTime(ms_kernel, E->ev0, E->ev1), FG_ERR_CHANNEL);
    return FG_OK;
}

extern "C" fg_status fg_film_composite(fg_engine *E, const fg_film_options *o, uint32_t W, uint32_t H,
                                       const uint8_t *gray, const float *h_eff, const float *tiles,
                                       int64_t n_tiles, uint8_t *out, float *d_norm, float *d_high, float *grain,
                                       fg_composite_stats *stats)
{
    std::string msg;
    if (!E || !gray || !h_eff || !tiles || !out) return fail(FG_ERR_INVALID_ARG, "fg_film_composite: null argument");
    if (!valid_options(o, msg) || !frame_ok(W, H, msg)) return fail(FG_ERR_INVALID_ARG, "fg_film_composite: " + msg);
    fg_film_plan plan;
    make_plan(*o, W, H, plan);
    const int64_t need = tile_values(W, H, plan.tile_step, plan.pad);
    if (n_tiles != need)
        return fail(FG_ERR_INVALID_ARG, "fg_film_composite: n_tile_values must be " + std::to_string(need));
    if (!kCompositeEnabled) return fail(FG_ERR_PIPELINE, std::string("fg_film_composite: ") + kCompositeOff);
    std::lock_guard<std::mutex> fl(E->film_mu);
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    const size_t n = (size_t)W * H;
    HIP_TRY(E->film_gray.ensure(n), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_heff.ensure(n * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->film_tiles.ensure((size_t)need * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemcpyAsync(E->film_gray.p, gray, n, hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpyAsync(E->film_heff.p, h_eff, n * 4, hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpyAsync(E->film_tiles.p, tiles, (size_t)need * 4, hipMemcpyHostToDevice, E->stream),
            FG_ERR_BUFFER_MAP);
    fg::FilmLayout L{};
    HIP_TRY(fg::film_layout_upload(W, H, plan.tile_step, plan.pad, E->film_layout, L, E->stream), FG_ERR_BUFFER_MAP);
    float ms = 0.0f;
    fg_status s = composite_device(E, plan, o->scanner_otf ? 1u : 0u, L, E->film_gray.as<uint8_t>(),
                                   E->film_heff.as<float>(), E->film_tiles.as<float>(), grain != nullptr, &ms);
    if (s != FG_OK) return s;
    float fs[4];
    HIP_TRY(hipMemcpy(fs, E->film_fstats.p, 16, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpy(out, E->film_out.p, n, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    if (d_norm) HIP_TRY(hipMemcpy(d_norm, E->film_dnorm.p, n * 4, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    if (d_high) HIP_TRY(hipMemcpy(d_high, E->film_dhigh.p, n * 4, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    if (grain) HIP_TRY(hipMemcpy(grain, E->film_grain.p, n * 4, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->high_mean = fs[0];
        stats->high_std = fs[1];
        stats->norm_factor = fs[2];
        stats->filter_sigma = plan.filter_sigma;
        stats->k_rad = plan.k_rad;
        stats->tiles_x = plan.tiles_x;
        stats->tiles_y = plan.tiles_y;
        stats->ms_kernels = ms;
    }
    return FG_OK;
}

extern "C" fg_status fg_render_film_bw(fg_engine *E, const uint8_t *pixels, uint32_t W, uint32_t H,
                                       uint32_t channels, const fg_film_options *o, uint8_t *out,
                                       fg_film_report *report)
{
    std::string msg;
    if (!E || !pixels || !out) return fail(FG_ERR_INVALID_ARG, "fg_render_film_bw: null argument");
    if (channels != 1 && channels != 3) return fail(FG_ERR_INVALID_ARG, "fg_render_film_bw: channels must be 1 or 3");
    if (!valid_options(o, msg) || !frame_ok(W, H, msg)) return fail(FG_ERR_INVALID_ARG, "fg_render_film_bw: " + msg);
    if (!kCompositeEnabled) return fail(FG_ERR_PIPELINE, std::string("fg_render_film_bw: ") + kCompositeOff);
    const auto t0 = Clock::now();
    fg_film_plan plan;
    make_plan(*o, W, H, plan);
    const std::string preset = lower(o->preset);
    const size_t n = (size_t)W * H;
    std::lock_guard<std::mutex> fl(E->film_mu);
    std::vector<float> heff(n);
    float ms_optics = 0.0f;
    {
        std::lock_guard<std::mutex> lk(E->mu);
        HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
        HIP_TRY(hipEventRecord(E->ev2, E->stream), FG_ERR_CHANNEL);
        fg_status s = upload_gray(E, pixels, W, H, channels, preset);
        if (s != FG_OK) return s;
        HIP_TRY(E->film_heff.ensure(n * 4), FG_ERR_BUFFER_RANGE);
        HIP_TRY(E->film_bbox.ensure(fg::film_bbox_size(W, H) * 4), FG_ERR_BUFFER_RANGE);
        HIP_TRY(fg::launch_film_optics(E->film_gray.as<uint8_t>(), W, H, E->film_bbox.as<int32_t>(),
                                       E->film_heff.as<float>(), E->stream),
                FG_ERR_PIPELINE);
        HIP_TRY(hipEventRecord(E->ev3, E->stream), FG_ERR_CHANNEL);
        HIP_TRY(hipMemcpyAsync(heff.data(), E->film_heff.p, n * 4, hipMemcpyDeviceToHost, E->stream),
                FG_ERR_BUFFER_MAP);
        HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
        (void)hipEventElapsedTime(&ms_optics, E->ev2, E->ev3);
    }
    const auto t1 = Clock::now();
    // one batched frame of the overlapping tiles, each rendered as render_tile_all_gpu
    // with the preset's parameters (image_pipeline.rs:436-515, :924-1014)
    fg_tile_params tp;
    fg_status s = fg_tile_params_init(&tp, o->preset, plan.grain_size_px, plan.tile_step, plan.tile_step,
                                      o->exposure, o->push, 0);
    if (s != FG_OK) return s;
    fg_frame_desc d{};
    d.width = W;
    d.height = H;
    d.tile = plan.tile_step;
    d.overlap = plan.pad;
    d.threads = o->threads;
    d.pixel_size = plan.pixel_size_um;
    d.emulsion = tile_emulsion_config(tp);
    d.exposure.photons_per_sq_unit = tp.photons_per_sq_unit;
    d.exposure.illumination_angle_rad = tp.illumination_angle_rad;
    d.exposure.crystal_absorption_coeff = tp.crystal_absorption_coeff;
    d.exposure.max_bounces = tp.max_bounces;
    d.exposure.photons_per_pixel = 0;
    d.development = fg_development_config{tp.threshold_photons, 0u, tp.filament_expansion,
                                          tp.clumping_proximity_factor, 0.85 + 0.05 * plan.push};
    d.scan = fg_scan_config{0u, 0u, tp.samples_per_pixel, FG_SCAN_DNEG, tp.silver_opacity, 1u, 0u};
    d.seed = o->seed;
    d.exposure_map = heff.data();
    d.density_mult = nullptr;
    d.synthesis = o->synthesis;
    fg_frame *F = nullptr;
    s = fg_frame_create(E, &d, &F);
    if (s != FG_OK) return s;
    struct FrameGuard {
        fg_frame *f;
        ~FrameGuard() { fg_frame_destroy(f); }
    } guard{F};
    const auto t2 = Clock::now();
    fg_frame_stats fst{};
    s = fg_frame_render(E, F, o->frame_mode, &fst);
    if (s != FG_OK) return s;
    const auto t3 = Clock::now();
    float ms_comp = 0.0f, fs[4];
    {
        std::lock_guard<std::mutex> lk(E->mu);
        HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
        s = composite_device(E, plan, o->scanner_otf ? 1u : 0u, *frame_layout(F), E->film_gray.as<uint8_t>(),
                             E->film_heff.as<float>(), frame_tile_dneg(F), false, &ms_comp);
        if (s != FG_OK) return s;
        HIP_TRY(hipMemcpy(fs, E->film_fstats.p, 16, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
        HIP_TRY(hipMemcpy(out, E->film_out.p, n, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    }
    const auto t4 = Clock::now();
    if (report) {
        memset(report, 0, sizeof *report);
        report->in_w = report->out_w = W;
        report->in_h = report->out_h = H;
        report->total_tiles = fst.tiles;
        report->total_crystals = fst.crystals;
        report->developed_crystals = fst.developed;
        report->total_clumps = fst.clumps;
        report->photons = fst.photons;
        report->rays = fst.rays;
        report->elapsed_seconds = std::chrono::duration<double>(t4 - t0).count();
        report->high_mean = fs[0];
        report->high_std = fs[1];
        report->norm_factor = fs[2];
        report->ms_luma_optics = ms_between(t0, t1);
        report->ms_synthesis = ms_between(t1, t2);
        report->ms_render = ms_between(t2, t3);
        report->ms_composite = ms_between(t3, t4);
        report->ms_kernel_optics = ms_optics;
        report->ms_kernel_composite = ms_comp;
        report->ms_kernel_transport = fst.ms_kernel_transport;
        report->ms_kernel_scan = fst.ms_kernel_scan;
    }
    return FG_OK;
}
