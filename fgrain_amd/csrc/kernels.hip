// gfx950 kernels of the photon-transport + scan path.
//
// Traversal everywhere is the stackless threaded walk of fg_layout.h: one thread
// follows i -> i + 1 on an interior hit and i -> skip otherwise, which visits nodes in
// the reference's depth-first order (bvh.rs:300-347, right child first) and therefore
// reports hits in the reference's callback order, without a per-lane stack in LDS or
// scratch.  Nodes are read as two 16-byte loads, crystal records as three.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace fg {

namespace {

constexpr float kTwoPi = 6.283185307179586f;

struct NodeView {
    float lo[3], hi[3];
    uint32_t skip, prim;
};

__device__ __forceinline__ NodeView load_node(const Node *nodes, uint32_t i)
{
    const float4 *p = reinterpret_cast<const float4 *>(nodes + i);
    const float4 a = p[0], b = p[1];
    NodeView v;
    v.lo[0] = a.x; v.lo[1] = a.y; v.lo[2] = a.z;
    v.hi[0] = a.w; v.hi[1] = b.x; v.hi[2] = b.y;
    v.skip = __float_as_uint(b.z);
    v.prim = __float_as_uint(b.w);
    return v;
}

struct Rec {
    float c[3], sens, q[4], h[3];
    uint32_t morph;
};

__device__ __forceinline__ Rec load_rec(const float4 *recs, uint32_t k)
{
    const float4 a = recs[3 * k], b = recs[3 * k + 1], c = recs[3 * k + 2];
    Rec r;
    r.c[0] = a.x; r.c[1] = a.y; r.c[2] = a.z; r.sens = a.w;
    r.q[0] = b.x; r.q[1] = b.y; r.q[2] = b.z; r.q[3] = b.w;
    r.h[0] = c.x; r.h[1] = c.y; r.h[2] = c.z;
    r.morph = __float_as_uint(c.w);
    return r;
}

// Sequential photon stream: draw k is word k & 3 of Philox block k >> 2.
struct PhotonRng {
    uint32_t p0, p1, k0, k1, blk, rem;
    uint32_t b0, b1, b2, b3;
    __device__ __forceinline__ void init(uint64_t photon, uint32_t key0, uint32_t key1)
    {
        p0 = (uint32_t)photon;
        p1 = (uint32_t)(photon >> 32);
        k0 = key0;
        k1 = key1;
        blk = 0;
        rem = 0;
    }
    __device__ __forceinline__ uint32_t next()
    {
        if (rem == 0) {
            uint32_t c[4] = {p0, p1, blk, kTagPhoton};
            philox(c, k0, k1);
            b0 = c[0]; b1 = c[1]; b2 = c[2]; b3 = c[3];
            blk++;
            rem = 4;
        }
        const uint32_t w = b0;
        b0 = b1; b1 = b2; b2 = b3;
        rem--;
        return w;
    }
    __device__ __forceinline__ float u() { return u24(next()); }
};

__device__ __forceinline__ void normalize3(float &x, float &y, float &z)
{
    const float n = sqrtf(x * x + y * y + z * z);
    x = x / n;
    y = y / n;
    z = z / n;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------------- transport ---
// One thread per launched packet; the step order of transport.rs:73-185 in f32 with
// the photon-keyed Philox stream (draw order: x, y jitter, theta, phi, Fresnel, then
// per bounce free path, one Bernoulli per absorbing hit, HG pair or halation draws,
// Russian roulette).
__global__ void __launch_bounds__(256) transport_kernel(TransportArgs A)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long n_b = 0, n_nodes = 0, n_prims = 0, n_abs = 0, n_tot = 0;
    if (p < A.n_photons) {
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
        float w = 1.0f;
        if (A.emap) {
            const float uu = fminf(fmaxf(x / A.W, 0.0f), 1.0f);
            const float vv = fminf(fmaxf(y / A.H, 0.0f), 1.0f);
            const uint32_t ix = min((uint32_t)roundf(uu * (float)(A.tw - 1)), A.tw - 1);
            const uint32_t iy = min((uint32_t)roundf(vv * (float)(A.th - 1)), A.th - 1);
            w = A.emap[(size_t)iy * A.estride + ix];
        }
        w = fmaxf(w, 0.0f);
        bool alive = w > 0.0f;
        float dx = 0.f, dy = 0.f, dz = 1.f;
        if (alive) {
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
            if (rng.u() < fr) alive = false;
        }
        float ox = x, oy = y, oz = 0.001f;
        for (uint32_t b = 0; alive && b < A.max_bounces; b++) {
            n_b++;
            float s;
            if (A.mu_t <= 0.0f) s = 1e6f;
            else s = -__logf(1.0f - rng.u()) / A.mu_t;
            const float step = fminf(s, A.step_cap);
            const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
            uint32_t i = 0;
            while (i < A.n_nodes) {
                const NodeView nd = load_node(A.nodes, i);
                n_nodes++;
                const bool hit = box_hit(nd.lo, nd.hi, ox, oy, oz, ix, iy, iz);
                if (hit && nd.prim != kInterior) {
                    const uint32_t first = nd.prim >> 4, cnt = nd.prim & 15u;
                    for (uint32_t k = first; k < first + cnt; k++) {
                        n_prims++;
                        const Rec r = load_rec(A.recs, k);
                        float t0, t1;
                        if (!crystal_hit(r.c, r.q, r.h, r.morph, ox, oy, oz, dx, dy, dz, t0, t1))
                            continue;
                        if (!(t0 < step)) continue;
                        const float L = fmaxf(fminf(t1, step) - t0, 0.0f);
                        if (!(L > 0.0f)) continue;
                        const float pa = 1.0f - __expf(-A.mu_abs * L);
                        const float lam = w * pa * r.sens * 5.0f;
                        const float fl = fminf(floorf(lam), 4294967040.0f);
                        const uint32_t whole = fl > 0.0f ? (uint32_t)fl : 0u;
                        const float frac = fminf(fmaxf(lam - (float)whole, 0.0f), 1.0f);
                        const uint32_t cntk = whole + (rng.u() < frac ? 1u : 0u);
                        if (cntk > 0u) {
                            atomicAdd(&A.counts[k], cntk);
                            n_abs++;
                            n_tot += cntk;
                        }
                        w *= fmaxf(1.0f - pa * 0.9f, 0.05f);
                    }
                }
                i = (hit && nd.prim == kInterior) ? i + 1 : nd.skip;
            }
            const float qx = ox + dx * step, qy = oy + dy * step, qz = oz + dz * step;
            if (qx < 0.0f || qx >= A.W || qy < 0.0f || qy >= A.H) break;
            if (qz <= 0.0f) break;
            if (qz >= A.D) {
                ox = qx;
                oy = qy;
                oz = A.D - 0.001f;
                // anti-halation backing (optics.rs:160-196)
                w *= A.hal_eff;
                if (w < 1e-4f) {
                    if (rng.u() < 0.1f) w /= 0.1f;
                    else break;
                }
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
                // a packet of zero weight deposits nothing from here on
                if (!(w > 0.0f)) break;
                continue;
            }
            ox = qx;
            oy = qy;
            oz = qz;
            // Henyey-Greenstein (optics.rs:113-143)
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
            if (w < 1e-3f) {
                if (rng.u() < 0.1f) w /= 0.1f;
                else break;
            }
        }
    }
    if (A.stats) {
        n_b = wave_sum(n_b);
        n_nodes = wave_sum(n_nodes);
        n_prims = wave_sum(n_prims);
        n_abs = wave_sum(n_abs);
        n_tot = wave_sum(n_tot);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&A.stats[0], n_b);
            atomicAdd(&A.stats[1], n_nodes);
            atomicAdd(&A.stats[2], n_prims);
            atomicAdd(&A.stats[3], n_abs);
            atomicAdd(&A.stats[4], n_tot);
        }
    }
}

// ------------------------------------------------------------------------ scan ---
// One thread per output pixel: ss x ss apertures of g x g stratified rays from
// z = D + 0.1 along -z (transmittance.rs:110-142).  For a ray along -z the
// reference's slab test (with +inf inverse components and NaN-ignoring min/max) is
// exactly "strictly inside the node's x-y rectangle" once the box reaches below the
// origin, which every box does; with float boxes rounded outward from the double
// boxes and float ray origins that comparison is exact, so the same crystals are
// tested as in the double-precision reference.
__global__ void __launch_bounds__(256) scan_kernel(ScanArgs A)
{
    const uint32_t px = blockIdx.x * 16 + (threadIdx.x & 15);
    const uint32_t py = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (px >= A.w_px || py >= A.h_px) return;
    const uint32_t pix = py * A.w_px + px;
    float acc = 0.0f;
    for (uint32_t ay = 0; ay < A.ss; ay++)
        for (uint32_t ax = 0; ax < A.ss; ax++)
            for (uint32_t sy = 0; sy < A.g; sy++)
                for (uint32_t sx = 0; sx < A.g; sx++) {
                    const uint32_t ray = ((ay * A.ss + ax) * A.g + sy) * A.g + sx;
                    uint32_t c[4] = {pix, ray >> 1, kTagScan, 0u};
                    philox(c, A.key0, A.key1);
                    const uint32_t wx = (ray & 1u) ? c[2] : c[0];
                    const uint32_t wy = (ray & 1u) ? c[3] : c[1];
                    const float subx = ((float)sx + u24(wx)) / (float)A.g;
                    const float suby = ((float)sy + u24(wy)) / (float)A.g;
                    const float apx = ((float)ax + subx) / (float)A.ss;
                    const float apy = ((float)ay + suby) / (float)A.ss;
                    const float ox = ((float)px + apx) * A.pw;
                    const float oy = ((float)py + apy) * A.ph;
                    float P = 0.0f;
                    uint32_t i = 0;
                    while (i < A.n_nodes) {
                        const NodeView nd = load_node(A.nodes, i);
                        const bool hit = nd.lo[0] < ox && ox < nd.hi[0] && nd.lo[1] < oy && oy < nd.hi[1];
                        if (hit && nd.prim != kInterior) {
                            const uint32_t first = nd.prim >> 4, cnt = nd.prim & 15u;
                            for (uint32_t k = first; k < first + cnt; k++) {
                                const Rec r = load_rec(A.recs, k);
                                float t0, t1;
                                if (crystal_hit(r.c, r.q, r.h, r.morph, ox, oy, A.z0, 0.0f, 0.0f, -1.0f, t0, t1))
                                    P += fmaxf(t1 - t0, 0.0f);
                            }
                        }
                        i = (hit && nd.prim == kInterior) ? i + 1 : nd.skip;
                    }
                    acc += expf(-A.sigma * P);
                }
    const float T = acc / (float)(A.g * A.g * A.ss * A.ss);
    float v;
    if (A.mode == 0u) {
        v = T;
    } else if (A.mode == 1u) {
        v = 1.0f - T;
    } else if (A.mode == 2u) {
        const float t = fmaxf(T, 1e-4f);
        v = fminf(fmaxf(-log10f(t) / 3.0f, 0.0f), 1.0f);
    } else {
        const double t = fmin(fmax((double)T, 1e-4), 1.0);
        v = (float)(-log10(t) + 0.05);
    }
    A.out[(size_t)py * A.ostride + px] = v;
}

// ------------------------------------------------------- traversal for tests ---
__global__ void __launch_bounds__(256) all_hits_kernel(const Node *nodes, uint32_t n_nodes,
                                                       const float4 *recs, const float *o,
                                                       const float *d, int64_t n_rays,
                                                       int32_t max_hits, int32_t *n_hits,
                                                       int32_t *slot, float *t0o, float *t1o)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rays) return;
    const float ox = o[3 * r], oy = o[3 * r + 1], oz = o[3 * r + 2];
    const float dx = d[3 * r], dy = d[3 * r + 1], dz = d[3 * r + 2];
    const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
    int32_t n = 0;
    uint32_t i = 0;
    while (i < n_nodes) {
        const NodeView nd = load_node(nodes, i);
        const bool hit = box_hit(nd.lo, nd.hi, ox, oy, oz, ix, iy, iz);
        if (hit && nd.prim != kInterior) {
            const uint32_t first = nd.prim >> 4, cnt = nd.prim & 15u;
            for (uint32_t k = first; k < first + cnt; k++) {
                const Rec rc = load_rec(recs, k);
                float t0, t1;
                if (crystal_hit(rc.c, rc.q, rc.h, rc.morph, ox, oy, oz, dx, dy, dz, t0, t1)) {
                    if (n < max_hits) {
                        slot[r * max_hits + n] = (int32_t)k;
                        t0o[r * max_hits + n] = t0;
                        t1o[r * max_hits + n] = t1;
                    }
                    n++;
                }
            }
        }
        i = (hit && nd.prim == kInterior) ? i + 1 : nd.skip;
    }
    n_hits[r] = n;
}

__global__ void __launch_bounds__(256) first_hit_kernel(const Node *nodes, uint32_t n_nodes,
                                                        const float4 *recs, const float *o,
                                                        const float *d, int64_t n_rays,
                                                        int32_t *slot, float *t0o, float *t1o)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rays) return;
    const float ox = o[3 * r], oy = o[3 * r + 1], oz = o[3 * r + 2];
    const float dx = d[3 * r], dy = d[3 * r + 1], dz = d[3 * r + 2];
    const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
    int32_t best = -1;
    float b0 = 0.0f, b1 = 0.0f;
    uint32_t i = 0;
    while (i < n_nodes) {
        const NodeView nd = load_node(nodes, i);
        const bool hit = box_hit(nd.lo, nd.hi, ox, oy, oz, ix, iy, iz);
        if (hit && nd.prim != kInterior) {
            const uint32_t first = nd.prim >> 4, cnt = nd.prim & 15u;
            for (uint32_t k = first; k < first + cnt; k++) {
                const Rec rc = load_rec(recs, k);
                float t0, t1;
                if (crystal_hit(rc.c, rc.q, rc.h, rc.morph, ox, oy, oz, dx, dy, dz, t0, t1) &&
                    (best < 0 || t0 < b0)) {
                    best = (int32_t)k;
                    b0 = t0;
                    b1 = t1;
                }
            }
        }
        i = (hit && nd.prim == kInterior) ? i + 1 : nd.skip;
    }
    slot[r] = best;
    t0o[r] = b0;
    t1o[r] = b1;
}

// ------------------------------------------------------------------- OTF ---------
// Separable [0.035, 0.93, 0.035] with clamp-to-edge (image_pipeline.rs:822-855),
// a 16 x 16 tile with a one-pixel halo staged in LDS, horizontal then vertical pass.
__global__ void __launch_bounds__(256) otf3_kernel(const float *in, float *out, uint32_t w,
                                                   uint32_t h)
{
    __shared__ float src[18][19];
    __shared__ float tmp[18][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x0 = (int)blockIdx.x * 16, y0 = (int)blockIdx.y * 16;
    for (int k = threadIdx.x; k < 18 * 18; k += 256) {
        const int ly = k / 18, lx = k % 18;
        const int gx = min(max(x0 + lx - 1, 0), (int)w - 1);
        const int gy = min(max(y0 + ly - 1, 0), (int)h - 1);
        src[ly][lx] = in[(size_t)gy * w + gx];
    }
    __syncthreads();
    const float k0 = 0.035f, k1 = 0.93f, k2 = 0.035f;
    for (int k = threadIdx.x; k < 18 * 16; k += 256) {
        const int ly = k / 16, lx = k % 16;
        const int gx = x0 + lx;
        // clamp-to-edge in x: at the frame border the neighbour is the pixel itself
        const float c = src[ly][lx + 1];
        const float l = (gx == 0) ? c : src[ly][lx];
        const float r = (gx + 1 >= (int)w) ? c : src[ly][lx + 2];
        tmp[ly][lx] = l * k0 + c * k1 + r * k2;
    }
    __syncthreads();
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx >= (int)w || gy >= (int)h) return;
    const float c = tmp[ty + 1][tx];
    const float u = (gy == 0) ? c : tmp[ty][tx];
    const float dn = (gy + 1 >= (int)h) ? c : tmp[ty + 2][tx];
    out[(size_t)gy * w + gx] = u * k0 + c * k1 + dn * k2;
}

}  // namespace

hipError_t launch_transport(const TransportArgs &a, hipStream_t s)
{
    if (a.n_photons == 0) return hipSuccess;
    const uint64_t blocks = (a.n_photons + 255) / 256;
    hipLaunchKernelGGL(transport_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_scan(const ScanArgs &a, hipStream_t s)
{
    if (a.w_px == 0 || a.h_px == 0) return hipSuccess;
    hipLaunchKernelGGL(scan_kernel, dim3((a.w_px + 15) / 16, (a.h_px + 15) / 16), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_all_hits(const Node *nodes, uint32_t n_nodes, const float4 *recs,
                           const float *o, const float *d, int64_t n_rays, int32_t max_hits,
                           int32_t *n_hits, int32_t *slot, float *t0, float *t1, hipStream_t s)
{
    if (n_rays == 0) return hipSuccess;
    hipLaunchKernelGGL(all_hits_kernel, dim3((uint32_t)((n_rays + 255) / 256)), dim3(256), 0, s,
                       nodes, n_nodes, recs, o, d, n_rays, max_hits, n_hits, slot, t0, t1);
    return hipGetLastError();
}

hipError_t launch_first_hit(const Node *nodes, uint32_t n_nodes, const float4 *recs,
                            const float *o, const float *d, int64_t n_rays, int32_t *slot,
                            float *t0, float *t1, hipStream_t s)
{
    if (n_rays == 0) return hipSuccess;
    hipLaunchKernelGGL(first_hit_kernel, dim3((uint32_t)((n_rays + 255) / 256)), dim3(256), 0, s,
                       nodes, n_nodes, recs, o, d, n_rays, slot, t0, t1);
    return hipGetLastError();
}

hipError_t launch_otf3(const float *in, float *out, uint32_t w, uint32_t h, hipStream_t s)
{
    if (w == 0 || h == 0) return hipSuccess;
    hipLaunchKernelGGL(otf3_kernel, dim3((w + 15) / 16, (h + 15) / 16), dim3(256), 0, s, in, out, w, h);
    return hipGetLastError();
}

}  // namespace fg
