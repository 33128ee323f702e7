// gfx950 kernels of the expected (differentiable) model: path replay of the photon
// streams, Poisson-tail development probability, expected transmittance product and
// the adjoint of each.
//
// Path replay: a packet is regenerated from its Philox counters (photon id and seed)
// in every pass instead of being stored.  Its draws are those of the transport kernel
// (transport.rs:73-185 order) minus the stochastic-rounding and Russian-roulette
// draws, which the expected model integrates out: a hit deposits its mean latent count
// 5 * E * w * pa * sens, and a packet ends when it leaves the slab, after max_bounces,
// or once its relative weight w and the weight's reflectance tangent both fall below a
// cutoff.  The weight is multiplicative in every parameter except g, so the gradient
// pass carries forward tangents of w with respect to mu_abs and the antihalation
// reflectance R along the path, and the score d/dg log p_HG(cos theta) summed over the
// scattering events so far for g.  All per-packet and per-ray arithmetic is float;
// sums over packets and rays are double atomics (their order does not change a float
// result beyond the last double bit).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "adjoint.h"
#include "device_common.h"

namespace fg {

namespace {

using namespace dev;

struct ReplayAcc {
    float e = 0.f, mu = 0.f, g = 0.f, r = 0.f;
    unsigned long long bounces = 0, events = 0;
};

template <bool kGrad>
__device__ __forceinline__ void replay_photon(const ReplayArgs &R, uint64_t p, ReplayAcc &acc)
{
    const TransportArgs &A = R.t;
    PhotonRng rng;
    rng.init(p, A.key0, A.key1);
    float x, y;
    if (A.ppp > 0) {
        const uint64_t pix = p / A.ppp;
        const uint32_t px = (uint32_t)(pix % A.tw), py = (uint32_t)(pix / A.tw);
        const float ux = rng.u(), uy = rng.u();
        x = ((float)px + ux) * A.pw;
        y = ((float)py + uy) * A.ph;
    } else {
        const float ux = rng.u(), uy = rng.u();
        x = ux * A.W;
        y = uy * A.H;
    }
    // nearest-pixel exposure lookup of the tile path (image_pipeline.rs:949-961)
    const float uu = fminf(fmaxf(x / A.W, 0.0f), 1.0f);
    const float vv = fminf(fmaxf(y / A.H, 0.0f), 1.0f);
    const uint32_t ex = min((uint32_t)roundf(uu * (float)(A.tw - 1)), A.tw - 1);
    const uint32_t ey = min((uint32_t)roundf(vv * (float)(A.th - 1)), A.th - 1);
    const size_t e_idx = (size_t)ey * A.estride + ex;
    const float E = fmaxf(A.emap[e_idx], 0.0f);
    if (!(E > 0.0f)) return;
    float dx, dy, dz;
    {
        const float th = rng.u() * A.theta_max;
        const float ph = rng.u() * kTwoPi;
        const float st = __sinf(th);
        dx = st * __cosf(ph);
        dy = st * __sinf(ph);
        dz = __cosf(th);
        normalize3(dx, dy, dz);
        const float cz = fminf(fmaxf(fabsf(dz), 0.0f), 1.0f);
        const float m = 1.0f - cz;
        const float fr = A.r0 + (1.0f - A.r0) * (m * m * m * m * m);
        if (rng.u() < fr) return;
    }
    float w = 1.0f, w_mu = 0.0f, w_r = 0.0f, score = 0.0f;
    float a_e = 0.0f, a_mu = 0.0f, a_g = 0.0f, a_r = 0.0f;
    float ox = x, oy = y, oz = 0.001f;
    const bool prune = A.reach >= 0.0f;
    for (uint32_t b = 0; b < A.max_bounces; b++) {
        acc.bounces++;
        float s;
        if (A.mu_t <= 0.0f) s = 1e6f;
        else s = -__logf(1.0f - rng.u()) / A.mu_t;
        const float step = fminf(s, A.step_cap);
        const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
        const float qx = ox + dx * step, qy = oy + dy * step, qz = oz + dz * step;
        const float mg = A.reach + 1e-3f + 1e-5f * (fabsf(ox) + fabsf(oy) + fabsf(oz) + step);
        const float slo0 = fminf(ox, qx) - mg, shi0 = fmaxf(ox, qx) + mg;
        const float slo1 = fminf(oy, qy) - mg, shi1 = fmaxf(oy, qy) + mg;
        const float slo2 = fminf(oz, qz) - mg, shi2 = fmaxf(oz, qz) + mg;
        uint32_t i = 0;
        while (i < A.n_nodes) {
            const NodeView nd = load_node(A.nodes, i);
            const bool near = !prune || (nd.lo[0] <= shi0 && nd.hi[0] >= slo0 && nd.lo[1] <= shi1 &&
                                         nd.hi[1] >= slo1 && nd.lo[2] <= shi2 && nd.hi[2] >= slo2);
            const bool hit = near && box_hit(nd.lo, nd.hi, ox, oy, oz, ix, iy, iz);
            if (hit && nd.prim != kInterior) {
                const uint32_t first = nd.prim >> 4, nprim = nd.prim & 15u;
                for (uint32_t k = first; k < first + nprim; k++) {
                    const Rec r = load_rec(A.recs, k);
                    float t0, t1;
                    if (!crystal_hit(r.c, r.q, r.h, r.morph, ox, oy, oz, dx, dy, dz, t0, t1)) continue;
                    if (!(t0 < step)) continue;
                    const float L = fmaxf(fminf(t1, step) - t0, 0.0f);
                    if (!(L > 0.0f)) continue;
                    acc.events++;
                    const float pa = 1.0f - __expf(-A.mu_abs * L);
                    if (kGrad) {
                        const float base = 5.0f * r.sens * R.gdose[k];
                        const float dep = w * pa;
                        a_e += base * dep;
                        a_mu += base * E * (w_mu * pa + w * L * (1.0f - pa));
                        a_r += base * E * (w_r * pa);
                        a_g += base * E * dep * score;
                    } else {
                        atomicAdd(&R.dose[k], (double)(((E * w) * pa) * r.sens * 5.0f));
                    }
                    const float m = 1.0f - pa * 0.9f;
                    const float dm = -0.9f * L * (1.0f - pa);
                    w_mu = w_mu * m + w * dm;
                    w_r = w_r * m;
                    w = w * m;
                }
            }
            i = (hit && nd.prim == kInterior) ? i + 1 : nd.skip;
        }
        if (qx < 0.0f || qx >= A.W || qy < 0.0f || qy >= A.H) break;
        if (qz <= 0.0f) break;
        if (qz >= A.D) {
            ox = qx;
            oy = qy;
            oz = A.D - 0.001f;
            const float inx = dx, iny = dy, inz = dz;
            // anti-halation backing (optics.rs:160-196) without its roulette
            w_r = w_r * A.hal_eff + w * R.hal_unit;
            w_mu = w_mu * A.hal_eff;
            w = w * A.hal_eff;
            if (rng.u() < 0.4f) {
                const float u1 = rng.u(), u2 = rng.u();
                const float rr = sqrtf(u1), phi = kTwoPi * u2;
                dx = rr * __cosf(phi);
                dy = rr * __sinf(phi);
                dz = -sqrtf(fmaxf(1.0f - u1, 0.0f));
                normalize3(dx, dy, dz);
            } else {
                dz = -fabsf(dz);
                normalize3(dx, dy, dz);
            }
            if (w < R.cut && w_r < R.cut) break;
            if (A.base > 0.0f) {
                if (!(inz > 1e-6f) || !(-dz > 1e-6f)) break;
                ox += A.base * (inx / inz + dx / (-dz));
                oy += A.base * (iny / inz + dy / (-dz));
                if (ox < 0.0f || ox >= A.W || oy < 0.0f || oy >= A.H) break;
            }
            continue;
        }
        ox = qx;
        oy = qy;
        oz = qz;
        {
            const float xi1 = rng.u(), xi2 = rng.u();
            float ct;
            if (fabsf(A.g) < 1e-4f) {
                ct = 2.0f * xi1 - 1.0f;
            } else {
                const float term = (1.0f - A.g * A.g) / (1.0f - A.g + 2.0f * A.g * xi1);
                ct = (1.0f + A.g * A.g - term * term) / (2.0f * A.g);
            }
            ct = fminf(fmaxf(ct, -1.0f), 1.0f);
            if (kGrad) score += hg_score(A.g, ct);
            const float stt = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
            const float phi = kTwoPi * xi2;
            float wx = dx, wy = dy, wz = dz;
            normalize3(wx, wy, wz);
            float ux, uy, uz;
            if (fabsf(wx) > 0.1f) { ux = wz; uy = 0.0f; uz = -wx; }
            else { ux = 0.0f; uy = -wz; uz = wy; }
            normalize3(ux, uy, uz);
            const float vx = wy * uz - wz * uy, vy = wz * ux - wx * uz, vz = wx * uy - wy * ux;
            const float a = stt * __cosf(phi), bb = stt * __sinf(phi);
            dx = (ux * a + vx * bb) + wx * ct;
            dy = (uy * a + vy * bb) + wy * ct;
            dz = (uz * a + vz * bb) + wz * ct;
            normalize3(dx, dy, dz);
        }
        w *= A.albedo;
        w_mu *= A.albedo;
        w_r *= A.albedo;
        if (w < R.cut && w_r < R.cut) break;
    }
    if (kGrad) {
        if (a_e != 0.0f) atomicAdd(&R.gmap[e_idx], (double)a_e);
        acc.mu += a_mu;
        acc.g += a_g;
        acc.r += a_r;
    }
}

__device__ __forceinline__ double wave_sum_d(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void flush_replay_stats(const ReplayAcc &acc, unsigned long long *stats)
{
    const unsigned long long b = wave_sum(acc.bounces), e = wave_sum(acc.events);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[1], b);
        atomicAdd(&stats[2], e);
    }
}

__global__ void __launch_bounds__(256) replay_dose_kernel(ReplayArgs R)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    ReplayAcc acc;
    if (p < R.t.n_photons) replay_photon<false>(R, p, acc);
    if (R.stats) flush_replay_stats(acc, R.stats);
}

// Gradient pass: parameter sums per block (fixed reduction order) into partials.
__global__ void __launch_bounds__(256) replay_grad_kernel(ReplayArgs R)
{
    __shared__ double red[4][3];
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    ReplayAcc acc;
    if (p < R.t.n_photons) replay_photon<true>(R, p, acc);
    if (R.stats) flush_replay_stats(acc, R.stats);
    const double m = wave_sum_d((double)acc.mu), g = wave_sum_d((double)acc.g),
                 r = wave_sum_d((double)acc.r);
    const uint32_t wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = m;
        red[wv][1] = g;
        red[wv][2] = r;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const uint32_t j = threadIdx.x;
        R.partials[4 * (size_t)blockIdx.x + j] = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
    }
    if (threadIdx.x == 3) R.partials[4 * (size_t)blockIdx.x + 3] = 0.0;
}

// Persistent replay (lanes take the next packet as soon as theirs ends; the path of
// every packet is that of replay_photon).  Block partial sums as replay_grad_kernel.
template <bool kPre>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 5))) replay_persistent_dose_kernel(ReplayArgs R, unsigned long long *work)
{
    LaneCounters cnt;
    transport_persistent<false, kReplayDose, kPre, kLeafQueue, kCandidates, true>(
        R.t, nullptr, 0u, nullptr, nullptr, nullptr, nullptr, work, R.t.n_photons, cnt, &R, nullptr, R.t.pre);
    if (R.stats) {
        const unsigned long long e = wave_sum((unsigned long long)cnt.abs);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&R.stats[1], cnt.bounces);
            atomicAdd(&R.stats[2], e);
            atomicAdd(&R.stats[3], cnt.atomics);
        }
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 5))) replay_persistent_grad_kernel(ReplayArgs R, unsigned long long *work)
{
    __shared__ double red[4][3];
    LaneCounters cnt;
    ReplaySums sums;
    transport_persistent<false, kReplayGrad, false, kLeafQueue, kCandidates, true>(
        R.t, nullptr, 0u, nullptr, nullptr, nullptr, nullptr, work, R.t.n_photons, cnt, &R, &sums, R.t.pre);
    if (R.stats) {
        const unsigned long long e = wave_sum((unsigned long long)cnt.abs);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&R.stats[1], cnt.bounces);
            atomicAdd(&R.stats[2], e);
        }
    }
    const double m = wave_sum_d((double)sums.mu), g = wave_sum_d((double)sums.g),
                 r = wave_sum_d((double)sums.r);
    const uint32_t wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = m;
        red[wv][1] = g;
        red[wv][2] = r;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const uint32_t j = threadIdx.x;
        R.partials[4 * (size_t)blockIdx.x + j] = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
    }
    if (threadIdx.x == 3) R.partials[4 * (size_t)blockIdx.x + 3] = 0.0;
}

// P[Poisson(lam) >= thr] and its derivative, the Poisson pmf at thr - 1.
__device__ __forceinline__ void poisson_tail(double lam, uint32_t thr, double &p, double &dp)
{
    if (thr == 0u) {
        p = 1.0;
        dp = 0.0;
        return;
    }
    if (!(lam > 0.0)) {
        p = 0.0;
        dp = thr == 1u ? 1.0 : 0.0;
        return;
    }
    const double lg = log(lam);
    dp = exp((double)(thr - 1u) * lg - lam - lgamma((double)thr));
    if (lam < (double)thr) {
        // tail series from k = thr; terms fall by lam / k < 1
        double t = exp((double)thr * lg - lam - lgamma((double)thr + 1.0));
        double sum = t;
        for (uint32_t k = thr + 1u; k < thr + 2000u && t > 1e-17 * sum; k++) {
            t *= lam / (double)k;
            sum += t;
        }
        p = sum;
    } else {
        double t = exp(-lam), head = t;
        for (uint32_t k = 1u; k < thr; k++) {
            t *= lam / (double)k;
            head += t;
        }
        p = 1.0 - head;
    }
    p = fmin(fmax(p, 0.0), 1.0);
}

__global__ void __launch_bounds__(256) dev_prob_kernel(const double *dose64, float *dose32,
                                                       const float *dose_in, float *prob,
                                                       float *dprob, uint32_t n, uint32_t thr)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float d;
    if (dose64) {
        d = (float)dose64[k];
        dose32[k] = d;
    } else {
        d = dose_in[k];
    }
    double p, dp;
    poisson_tail((double)d, thr, p, dp);
    prob[k] = (float)p;
    dprob[k] = (float)dp;
}

// Expected clump boost of development.rs:164-171: a clump of size s multiplies the
// expansion by 1.15 + 0.18 (min(s, 6) - 1), a lone crystal by 1.  The expected model
// takes s = 1 + K (K = expected number of developed touching neighbours) and joins
// 1 and the size-2 boost linearly for 1 <= s <= 2.
__device__ __forceinline__ void boost_of(float s, float &b, float &db)
{
    if (s <= 1.0f) { b = 1.0f; db = 0.0f; }
    else if (s <= 2.0f) { b = 1.0f + 0.33f * (s - 1.0f); db = 0.33f; }
    else if (s < 6.0f) { b = 1.15f + 0.18f * (s - 1.0f); db = 0.18f; }
    else { b = 2.05f; db = 0.0f; }
}

__global__ void __launch_bounds__(256) boost_kernel(const uint32_t *ptr, const uint32_t *idx,
                                                    const float *prob, float *boost, float *dboost,
                                                    uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float K = 0.0f;
    for (uint32_t e = ptr[i]; e < ptr[i + 1]; e++) K += prob[idx[e]];
    float b, db;
    boost_of(1.0f + K, b, db);
    boost[i] = b;
    dboost[i] = db;
}

__global__ void __launch_bounds__(256) boost_grad_kernel(const uint32_t *ptr, const uint32_t *idx,
                                                         const double *gboost, const float *dboost,
                                                         double *gprob, uint32_t n)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double acc = 0.0;
    for (uint32_t e = ptr[j]; e < ptr[j + 1]; e++) {
        const uint32_t i = idx[e];
        acc += gboost[i] * (double)dboost[i];
    }
    gprob[j] += acc;
}

// Transmittance product along the ray at (ox, oy) going -z: every crystal whose
// boosted shape the ray crosses contributes 1 - prob * (1 - exp(-sigma * chord)).
// Node tests are those of the scan (strict x-y containment on boxes that hold the
// largest boost).
template <bool kAdjoint>
__device__ __forceinline__ float ray_product(const ExpScanArgs &A, float ox, float oy, float coef)
{
    float prod = 1.0f;
    uint32_t i = 0;
    while (i < A.n_nodes) {
        const NodeView nd = load_node(A.nodes, i);
        const bool hit = nd.lo[0] < ox && ox < nd.hi[0] && nd.lo[1] < oy && oy < nd.hi[1];
        if (hit && nd.prim != kInterior) {
            const uint32_t first = nd.prim >> 4, cnt = nd.prim & 15u;
            for (uint32_t k = first; k < first + cnt; k++) {
                const float4 a = A.recs[3 * k], c = A.recs[3 * k + 2];
                const uint32_t slot = __float_as_uint(a.w);
                const float bs = A.boost[slot];
                const float4 cs = make_float4(c.x * bs, c.y * bs, c.z * bs, c.w);
                if (!sphere_may_hit_z(a, cs, ox, oy)) continue;
                const Rec r = load_rec(A.recs, k);
                float chord, dchord;
                if (!crystal_chord_scaled(r.c, r.q, r.h, r.morph, bs, ox, oy, A.z0, 0.0f, 0.0f, -1.0f,
                                          chord, dchord))
                    continue;
                const float ex = expf(-A.sigma * chord);
                const float am = 1.0f - ex;
                const float p = A.prob[slot];
                const float f = 1.0f - p * am;
                if (kAdjoint) {
                    // coef holds dL/dT_ray (mean weight included) times the full product;
                    // f > 0 since prob <= 1 and exp(-sigma chord) > 0
                    if (f > 0.0f) {
                        const float rest = coef / f;
                        atomicAdd(&A.gprob[slot], (double)(-am * rest));
                        atomicAdd(&A.gboost[slot], (double)(-p * A.sigma * ex * dchord * rest));
                    }
                } else {
                    prod *= f;
                }
            }
        }
        i = (hit && nd.prim == kInterior) ? i + 1 : nd.skip;
    }
    return prod;
}

// dvalue/dT of the scan mode map (scan_value)
__device__ __forceinline__ float scan_value_grad(float T, uint32_t mode)
{
    constexpr float kLn10 = 2.302585092994046f;
    if (mode == 0u) return 1.0f;
    if (mode == 1u) return -1.0f;
    if (mode == 2u) return T > 1e-3f ? -1.0f / (3.0f * T * kLn10) : 0.0f;
    return T > 1e-4f ? -1.0f / (T * kLn10) : 0.0f;
}

template <bool kAdjoint>
__global__ void __launch_bounds__(256) expected_scan_kernel(ExpScanArgs A)
{
    const uint32_t Rn = A.g * A.g * A.ss * A.ss;
    const uint32_t L = scan_lanes(Rn, A.rays_per_lane);
    const uint32_t pix = blockIdx.x * (256u / L) + threadIdx.x / L;
    const bool valid = pix < A.w_px * A.h_px;
    const uint32_t px = valid ? pix % A.w_px : 0u, py = valid ? pix / A.w_px : 0u;
    const uint32_t sub = threadIdx.x % L;
    // pass 1: mean product over the pixel's rays (fixed xor-tree over the L lanes)
    float acc = 0.0f;
    if (valid) {
        for (uint32_t r = sub; r < Rn; r += L) {
            const uint32_t sx = r % A.g, sy = (r / A.g) % A.g, ax = (r / (A.g * A.g)) % A.ss,
                           ay = r / (A.g * A.g * A.ss);
            float ox, oy;
            scan_origin(pix, px, py, ax, ay, sx, sy, A.g, A.ss, A.pw, A.ph, A.key0, A.key1, ox, oy);
            acc += ray_product<false>(A, ox, oy, 0.0f);
        }
    }
    for (uint32_t off = L >> 1; off >= 1u; off >>= 1) acc += __shfl_xor(acc, (int)off, 64);
    const float T = acc / (float)Rn;
    if (!kAdjoint) {
        if (valid && sub == 0u) A.out[(size_t)py * A.w_px + px] = scan_value(T, A.mode);
        return;
    }
    // pass 2: scatter dL/dprob along every ray of the pixel
    if (!valid) return;
    const float gT = A.gout[(size_t)py * A.w_px + px] * scan_value_grad(T, A.mode);
    if (gT == 0.0f) return;
    const float c = gT / (float)Rn;
    for (uint32_t r = sub; r < Rn; r += L) {
        const uint32_t sx = r % A.g, sy = (r / A.g) % A.g, ax = (r / (A.g * A.g)) % A.ss,
                       ay = r / (A.g * A.g * A.ss);
        float ox, oy;
        scan_origin(pix, px, py, ax, ay, sx, sy, A.g, A.ss, A.pw, A.ph, A.key0, A.key1, ox, oy);
        const float prod = ray_product<false>(A, ox, oy, 0.0f);
        ray_product<true>(A, ox, oy, c * prod);
    }
}

__global__ void __launch_bounds__(256) dose_grad_kernel(const double *gprob, const float *dprob,
                                                        float *gdose, uint32_t n)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) gdose[k] = (float)(gprob[k] * (double)dprob[k]);
}

__global__ void __launch_bounds__(256) reduce_partials_kernel(const double *partials, uint32_t blocks,
                                                              double *out)
{
    __shared__ double red[256][4];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < blocks; b += 256u)
        for (int j = 0; j < 4; j++) s[j] += partials[4 * (size_t)b + j];
    for (int j = 0; j < 4; j++) red[threadIdx.x][j] = s[j];
    __syncthreads();
    for (uint32_t h = 128u; h > 0u; h >>= 1) {
        if (threadIdx.x < h)
            for (int j = 0; j < 4; j++) red[threadIdx.x][j] += red[threadIdx.x + h][j];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = red[0][threadIdx.x];
}

__global__ void __launch_bounds__(256) f64_to_f32_kernel(const double *in, float *out, uint32_t n)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = (float)in[k];
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255u) / 256u); }

}  // namespace

int replay_blocks(uint64_t n_photons) { return (int)blocks_for(n_photons); }

hipError_t launch_replay_dose(const ReplayArgs &a, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (a.t.n_photons == 0) return hipSuccess;
    replay_dose_kernel<<<blocks_for(a.t.n_photons), 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_replay_grad(const ReplayArgs &a, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (a.t.n_photons == 0) return hipSuccess;
    replay_grad_kernel<<<blocks_for(a.t.n_photons), 256, 0, s>>>(a);
    return hipGetLastError();
}

int replay_persistent_blocks_per_cu()
{
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, reinterpret_cast<const void *>(replay_persistent_dose_kernel<true>),
                                                     256, 0) != hipSuccess)
        a = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void *>(replay_persistent_grad_kernel),
                                                     256, 0) != hipSuccess)
        b = 1;
    return std::max(1, std::min(a, b));
}

hipError_t launch_replay_persistent_dose(const ReplayArgs &a, unsigned long long *work, int grid, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (a.t.n_photons == 0) return hipSuccess;
    if (!a.t.pre) return hipErrorInvalidValue;   // the walk's pre-test reads these records
    // the f64 deposits are always pre-reduced across the wave: with them the dose pass of
    // the 512 x 512 adjoint tile takes 72 ms instead of 81 (the counts transport, whose
    // u32 atomics are cheaper, leaves it to the tuning)
    replay_persistent_dose_kernel<true><<<grid, 256, 0, s>>>(a, work);
    return hipGetLastError();
}

hipError_t launch_replay_persistent_grad(const ReplayArgs &a, unsigned long long *work, int grid, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (a.t.n_photons == 0) return hipSuccess;
    if (!a.t.pre) return hipErrorInvalidValue;   // the walk's pre-test reads these records
    replay_persistent_grad_kernel<<<grid, 256, 0, s>>>(a, work);
    return hipGetLastError();
}

hipError_t launch_dev_prob(const double *dose64, float *dose32, float *prob, float *dprob,
                           uint32_t n, uint32_t thr, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n == 0) return hipSuccess;
    dev_prob_kernel<<<blocks_for(n), 256, 0, s>>>(dose64, dose32, nullptr, prob, dprob, n, thr);
    return hipGetLastError();
}

hipError_t launch_dev_prob_from(const float *dose32, float *prob, float *dprob, uint32_t n,
                                uint32_t thr, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n == 0) return hipSuccess;
    dev_prob_kernel<<<blocks_for(n), 256, 0, s>>>(nullptr, nullptr, dose32, prob, dprob, n, thr);
    return hipGetLastError();
}

static uint32_t expected_scan_grid(const ExpScanArgs &a)
{
    const uint32_t L = scan_lanes(a.g * a.g * a.ss * a.ss, a.rays_per_lane);
    const uint64_t px = (uint64_t)a.w_px * a.h_px;
    return (uint32_t)((px + (256u / L) - 1u) / (256u / L));
}

hipError_t launch_expected_scan(const ExpScanArgs &a, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if ((uint64_t)a.w_px * a.h_px == 0) return hipSuccess;
    expected_scan_kernel<false><<<expected_scan_grid(a), 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_expected_scan_adjoint(const ExpScanArgs &a, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if ((uint64_t)a.w_px * a.h_px == 0) return hipSuccess;
    expected_scan_kernel<true><<<expected_scan_grid(a), 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_boost(const uint32_t *ptr, const uint32_t *idx, const float *prob, float *boost,
                        float *dboost, uint32_t n, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n == 0) return hipSuccess;
    boost_kernel<<<blocks_for(n), 256, 0, s>>>(ptr, idx, prob, boost, dboost, n);
    return hipGetLastError();
}

hipError_t launch_boost_grad(const uint32_t *ptr, const uint32_t *idx, const double *gboost,
                             const float *dboost, double *gprob, uint32_t n, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n == 0) return hipSuccess;
    boost_grad_kernel<<<blocks_for(n), 256, 0, s>>>(ptr, idx, gboost, dboost, gprob, n);
    return hipGetLastError();
}

hipError_t launch_dose_grad(const double *gprob, const float *dprob, float *gdose, uint32_t n,
                            hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n == 0) return hipSuccess;
    dose_grad_kernel<<<blocks_for(n), 256, 0, s>>>(gprob, dprob, gdose, n);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double *partials, uint32_t blocks, double *out, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    reduce_partials_kernel<<<1, 256, 0, s>>>(partials, blocks, out);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double *in, float *out, uint32_t n, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n == 0) return hipSuccess;
    f64_to_f32_kernel<<<blocks_for(n), 256, 0, s>>>(in, out, n);
    return hipGetLastError();
}

}  // namespace fg
