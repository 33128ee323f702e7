// Device functions shared by the per-tile kernels (kernels.hip) and the batched
// pipeline (batched.hip): node and record loads, the photon Philox stream, the
// transport of one packet and the scan of one pixel.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace fg {
namespace dev {

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

// Raw buffer loads for the transport's hot loops: a wave-uniform resource (base address
// in scalar registers; word 3 as for every gfx9 raw buffer) and a 32-bit byte offset per
// lane, so an address costs one VALU operation instead of a 64-bit shift and add
// (transport 1246 -> 1224 ms Tri-X, 3776 -> 3696 ms T-Max).  The host refuses node or
// record arrays of 4 GiB or more (offsets_fit_32).
typedef int BufWords __attribute__((ext_vector_type(4)));
constexpr int kBufferWord3 = 0x00020000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)0xFFFFFFFF, kBufferWord3);
}
__device__ __forceinline__ float4 buf_load4(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    const BufWords v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_float4(__int_as_float(v.x), __int_as_float(v.y), __int_as_float(v.z), __int_as_float(v.w));
}
__device__ __forceinline__ NodeView load_node_buf(__amdgpu_buffer_rsrc_t r, uint32_t i)
{
    const uint32_t o = i << 5;
    const float4 a = buf_load4(r, o), b = buf_load4(r, o + 16u);
    NodeView v;
    v.lo[0] = a.x; v.lo[1] = a.y; v.lo[2] = a.z;
    v.hi[0] = a.w; v.hi[1] = b.x; v.hi[2] = b.y;
    v.skip = __float_as_uint(b.z);
    v.prim = __float_as_uint(b.w);
    return v;
}
__device__ __forceinline__ Rec load_rec_buf(__amdgpu_buffer_rsrc_t r, uint32_t k)
{
    const uint32_t o = k * 48u;
    const float4 a = buf_load4(r, o), b = buf_load4(r, o + 16u), c = buf_load4(r, o + 32u);
    Rec q;
    q.c[0] = a.x; q.c[1] = a.y; q.c[2] = a.z; q.sens = a.w;
    q.q[0] = b.x; q.q[1] = b.y; q.q[2] = b.z; q.q[3] = b.w;
    q.h[0] = c.x; q.h[1] = c.y; q.h[2] = c.z;
    q.morph = __float_as_uint(c.w);
    return q;
}

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

// Conservative pre-test before crystal_hit for a ray along -z through (ox, oy): the
// crystal lies inside the sphere about its centre of radius |h| (box) or max(h)
// (ellipsoid); when the ray passes farther than that radius in the x-y plane (plus a
// margin well above the float rounding of crystal_hit) the exact test would miss too,
// so skipping it changes no result.  The rays of one pixel sit in adjacent lanes, so
// the test mostly rejects for the whole group at once.
// the same test with the bounding-sphere radius R = sqrtf(r2) precomputed (down_constants)
__device__ __forceinline__ bool sphere_may_hit_z_r(const float4 &a, float R, float ox, float oy)
{
    const float vx = a.x - ox, vy = a.y - oy;
    const float d2 = vx * vx + vy * vy;
    const float m = R * 1.001f + 1e-4f * (fabsf(vx) + fabsf(vy)) + 1e-4f;
    return d2 <= m * m;
}

__device__ __forceinline__ bool sphere_may_hit_z(const float4 &a, const float4 &c, float ox, float oy)
{
    const float vx = a.x - ox, vy = a.y - oy;
    const float d2 = vx * vx + vy * vy;
    const uint32_t morph = __float_as_uint(c.w);
    const float r2 = morph == 0u ? c.x * c.x + c.y * c.y + c.z * c.z
                                 : fmaxf(fmaxf(c.x, c.y), c.z) * fmaxf(fmaxf(c.x, c.y), c.z);
    const float m = sqrtf(r2) * 1.001f + 1e-4f * (fabsf(vx) + fabsf(vy)) + 1e-4f;
    return d2 <= m * m;
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

// Henyey-Greenstein score: d/dg log p_g(mu), p_g(mu) = (1 - g^2) / (2 (1 + g^2 - 2 g mu)^1.5)
__device__ __forceinline__ float hg_score(float g, float ct)
{
    return -2.0f * g / (1.0f - g * g) + 3.0f * (ct - g) / (1.0f + g * g - 2.0f * g * ct);
}

// Per-lane sums of a gradient replay: d/d(mu_abs), d/dg, d/dR.
struct ReplaySums {
    float mu = 0.f, g = 0.f, r = 0.f;
};

struct PhotonCounters {
    unsigned long long bounces = 0, nodes = 0, prims = 0, abs = 0, tot = 0;
};

// Counters of the persistent transport: bounces and node steps are counted per wave
// (popcounts of ballots, wave-uniform, so they live in scalar registers), crystal tests,
// absorbing hits and latent photons per lane in 32 bits (a lane handles far fewer than
// 2^32 of each).
// atomics: global atomic additions of deposits the lane issued (after the wave
// pre-reduction); packets: packets the wave started (wave-uniform), which must come to
// the launch's packet count.
struct LaneCounters {
    unsigned long long bounces = 0, nodes = 0, packets = 0;
    uint32_t prims = 0, abs = 0, tot = 0, atomics = 0;
};

__device__ __forceinline__ void flush_lane_counters(const LaneCounters &c, unsigned long long *stats)
{
    const unsigned long long pr = wave_sum((unsigned long long)c.prims),
                             ab = wave_sum((unsigned long long)c.abs),
                             t = wave_sum((unsigned long long)c.tot),
                             at = wave_sum((unsigned long long)c.atomics);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], c.bounces);
        atomicAdd(&stats[1], c.nodes);
        atomicAdd(&stats[2], pr);
        atomicAdd(&stats[3], ab);
        atomicAdd(&stats[4], t);
        atomicAdd(&stats[5], at);
        atomicAdd(&stats[7], c.packets);
    }
}

// Wave pre-reduction of deposits before the global atomics (record_absorption,
// crystal.rs:221-223, is a relaxed fetch_add per absorbing hit).  The lanes of a wave
// that deposit in the same step mostly follow neighbouring packets of one pixel, so they
// often deposit into the same crystal.  Groups are formed one crystal at a time: the
// lowest pending lane names a slot, a ballot finds every pending lane with that slot;
// a group of one lane issues its own atomic, a larger group adds its values into the
// wave's LDS word and the naming lane moves the sum to HBM with one atomic.  Integer
// sums do not depend on their order, so the counts are bit-identical to one atomic per
// lane.  Called by the active lanes of a wave (any subset); acc is the wave's LDS word
// (zero on entry and on return); returns the global atomics this lane issued.
template <class T>
__device__ __forceinline__ uint32_t wave_aggregated_add(T *dst, uint32_t slot, T val, bool want, T *acc)
{
    const uint64_t pend = __builtin_amdgcn_ballot_w64(want);
    if (pend == 0ull) return 0u;
    const uint32_t leader = (uint32_t)__builtin_ctzll(pend);
    const uint32_t key = __builtin_amdgcn_readlane(slot, leader);
    const bool mine = want && slot == key;
    const uint64_t grp = __builtin_amdgcn_ballot_w64(mine);
    uint32_t issued = 0;
    if (want && !mine) {
        atomicAdd(&dst[slot], val);
        issued = 1u;
    }
    if ((grp & (grp - 1)) == 0) {
        if (mine) {
            atomicAdd(&dst[key], val);
            issued = 1u;
        }
    } else {
        if (mine) atomicAdd(acc, val);
        __builtin_amdgcn_wave_barrier();
        const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
        if (lane == leader) {
            atomicAdd(&dst[key], atomicExch(acc, (T)0));
            issued = 1u;
        }
        __builtin_amdgcn_wave_barrier();
    }
    return issued;
}

__device__ __forceinline__ void flush_counters(PhotonCounters &c, unsigned long long *stats)
{
    const unsigned long long b = wave_sum(c.bounces), n = wave_sum(c.nodes),
                             pr = wave_sum(c.prims), ab = wave_sum(c.abs), t = wave_sum(c.tot);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], b);
        atomicAdd(&stats[1], n);
        atomicAdd(&stats[2], pr);
        atomicAdd(&stats[3], ab);
        atomicAdd(&stats[4], t);
    }
}

// One launched packet, the step order of transport.rs:73-185 in f32 with the
// photon-keyed Philox stream.  Draw order: x, y jitter, theta, phi, Fresnel; per bounce
// the free path, one Bernoulli per absorbing hit, then the HG pair or the halation
// draws, then Russian roulette.  Latent photons go to counts[leaf slot].
__device__ __forceinline__ void run_photon(const TransportArgs &A, uint64_t p, PhotonCounters &cnt)
{
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
        // nearest-pixel lookup of the tile path (image_pipeline.rs:949-961)
        const float uu = fminf(fmaxf(x / A.W, 0.0f), 1.0f);
        const float vv = fminf(fmaxf(y / A.H, 0.0f), 1.0f);
        const uint32_t ix = min((uint32_t)roundf(uu * (float)(A.tw - 1)), A.tw - 1);
        const uint32_t iy = min((uint32_t)roundf(vv * (float)(A.th - 1)), A.th - 1);
        w = A.emap[(size_t)iy * A.estride + ix];
    }
    w = fmaxf(w, 0.0f);
    if (!(w > 0.0f)) return;
    float dx, dy, dz;
    {
        const float th = rng.u() * A.theta_max;
        const float ph = rng.u() * kTwoPi;
        const float st = __sinf(th);
        dx = st * __cosf(ph);
        dy = st * __sinf(ph);
        dz = __cosf(th);
        normalize3(dx, dy, dz);
        // Schlick Fresnel rejection (optics.rs:147-154)
        const float cz = fminf(fmaxf(fabsf(dz), 0.0f), 1.0f);
        const float m = 1.0f - cz;
        const float fr = A.r0 + (1.0f - A.r0) * (m * m * m * m * m);
        if (rng.u() < fr) return;
    }
    float ox = x, oy = y, oz = 0.001f;
    for (uint32_t b = 0; b < A.max_bounces; b++) {
        cnt.bounces++;
        float s;
        if (A.mu_t <= 0.0f) s = 1e6f;
        else s = -__logf(1.0f - rng.u()) / A.mu_t;
        const float step = fminf(s, A.step_cap);
        const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
        // bounding box of the step segment grown by the protrusion reach and a margin
        // far above float rounding of the coordinates
        const bool prune = A.reach >= 0.0f;
        const float ex = ox + dx * step, ey = oy + dy * step, ez = oz + dz * step;
        const float mg = A.reach + 1e-3f + 1e-5f * (fabsf(ox) + fabsf(oy) + fabsf(oz) + step);
        const float slo0 = fminf(ox, ex) - mg, shi0 = fmaxf(ox, ex) + mg;
        const float slo1 = fminf(oy, ey) - mg, shi1 = fmaxf(oy, ey) + mg;
        const float slo2 = fminf(oz, ez) - mg, shi2 = fmaxf(oz, ez) + mg;
        uint32_t i = 0;
        while (i < A.n_nodes) {
            const NodeView nd = load_node(A.nodes, i);
            cnt.nodes++;
            const bool near = !prune || (nd.lo[0] <= shi0 && nd.hi[0] >= slo0 && nd.lo[1] <= shi1 &&
                                         nd.hi[1] >= slo1 && nd.lo[2] <= shi2 && nd.hi[2] >= slo2);
            const bool hit = near && box_hit(nd.lo, nd.hi, ox, oy, oz, ix, iy, iz);
            if (hit && nd.prim != kInterior) {
                const uint32_t first = nd.prim >> 4, nprim = nd.prim & 15u;
                for (uint32_t k = first; k < first + nprim; k++) {
                    cnt.prims++;
                    const Rec r = load_rec(A.recs, k);
                    float t0, t1;
                    if (!crystal_hit(r.c, r.q, r.h, r.morph, ox, oy, oz, dx, dy, dz, t0, t1)) continue;
                    if (!(t0 < step)) continue;
                    const float L = fmaxf(fminf(t1, step) - t0, 0.0f);
                    if (!(L > 0.0f)) continue;
                    // Beer-Lambert + stochastic rounding (transport.rs:117-134)
                    const float pa = 1.0f - __expf(-A.mu_abs * L);
                    const float lam = w * pa * r.sens * 5.0f;
                    const float fl = fminf(floorf(lam), 4294967040.0f);
                    const uint32_t whole = fl > 0.0f ? (uint32_t)fl : 0u;
                    const float frac = fminf(fmaxf(lam - (float)whole, 0.0f), 1.0f);
                    const uint32_t k_lat = whole + (rng.u() < frac ? 1u : 0u);
                    if (k_lat > 0u) {
                        atomicAdd(&A.counts[k], k_lat);
                        cnt.abs++;
                        cnt.tot += k_lat;
                    }
                    w *= fmaxf(1.0f - pa * 0.9f, 0.05f);
                }
            }
            i = (hit && nd.prim == kInterior) ? i + 1 : nd.skip;
        }
        const float qx = ox + dx * step, qy = oy + dy * step, qz = oz + dz * step;
        if (qx < 0.0f || qx >= A.W || qy < 0.0f || qy >= A.H) return;
        if (qz <= 0.0f) return;
        if (qz >= A.D) {
            ox = qx;
            oy = qy;
            oz = A.D - 0.001f;
            const float inx = dx, iny = dy, inz = dz;
            // anti-halation backing (optics.rs:160-196)
            w *= A.hal_eff;
            if (w < 1e-4f) {
                if (rng.u() < 0.1f) w /= 0.1f;
                else return;
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
            if (!(w > 0.0f)) return;
            if (A.base > 0.0f) {
                if (!(inz > 1e-6f) || !(-dz > 1e-6f)) return;
                ox += A.base * (inx / inz + dx / (-dz));
                oy += A.base * (iny / inz + dy / (-dz));
                if (ox < 0.0f || ox >= A.W || oy < 0.0f || oy >= A.H) return;
            }
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
            else return;
        }
    }
}

// ---------------------------------------------------------------------------------
// Persistent transport.  Each lane runs a small state machine (wait for a packet,
// start a bounce, walk the BVH, process a pending leaf, finish the bounce) so that a
// lane whose packet ends takes the next packet immediately instead of idling until
// the slowest lane of its wave finishes.  Packets are handed out from a wave-local
// pool refilled with one atomic per `chunk` packets, consecutive ids to consecutive
// lanes (the ids of one pixel).  Node steps continue until `wait_pct` percent of the
// walking lanes hold a full leaf queue (or have ended their walk), then the queued
// leaves are processed together.  For every packet the draws, the order of its hits and
// its deposits are those of run_photon.

// Per-lane tile state kept through the walk: offsets into the frame's node, record and
// count arrays, the node count and the patch extents.  The constants a packet needs only
// when it starts (keys, pixel grid, exposure placement) are read then (PacketStart).
struct LaneTile {
    uint32_t node_base, crystal_base, n_nodes;
    float W, H;
};

struct PacketStart {
    const float *emap;
    uint32_t tw, th, key0, key1;
    float pw, ph;
};

__device__ __forceinline__ void lane_tile_single(const TransportArgs &C, LaneTile &T, PacketStart &S)
{
    T.node_base = 0u;
    T.crystal_base = 0u;
    T.n_nodes = C.n_nodes;
    T.W = C.W;
    T.H = C.H;
    S.emap = C.emap;
    S.tw = C.tw;
    S.th = C.th;
    S.key0 = C.key0;
    S.key1 = C.key1;
    S.pw = C.pw;
    S.ph = C.ph;
}

// index of the tile of frame packet id g (last tile with photon_base <= g)
__device__ __forceinline__ uint32_t tile_of_packet(const TileDev *tiles, uint32_t n_tiles, uint64_t g)
{
    uint32_t lo = 0, hi = n_tiles;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tiles[mid].photon_base <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void lane_tile_fill(const TileDev &t, const float *emap, uint32_t estride, LaneTile &T,
                                               PacketStart &S)
{
    T.node_base = t.node_base;
    T.crystal_base = t.crystal_base;
    T.n_nodes = t.n_nodes;
    T.W = t.W;
    T.H = t.H;
    S.emap = emap + (size_t)t.y0 * estride + t.x0;
    S.tw = t.tw;
    S.th = t.th;
    S.key0 = t.key0;
    S.key1 = t.key1;
    S.pw = t.pw;
    S.ph = t.ph;
}

enum : uint32_t { kNeed = 0, kStart = 1, kWalk = 2, kDone = 3 };

// The node test of the transport walk as one floating-point comparison: the node box
// overlaps the segment box [s, h] (lo <= h and hi >= s on every axis) and box_hit holds
// (tx >= te and tx > 0).  With denormals kept, a - b <= 0 exactly when a <= b for finite
// a and b (and for the unbounded segment box of an unpruned walk), and tx >= max(te,
// smallest denormal) exactly when tx >= te and tx > 0, so the maximum of the seven
// differences is <= 0 exactly when both tests hold.
__device__ __forceinline__ bool near_box_hit(const NodeView &nd, float s0, float s1, float s2, float h0, float h1,
                                             float h2, float ox, float oy, float oz, float ix, float iy, float iz)
{
    const float t1x = (nd.lo[0] - ox) * ix, t2x = (nd.hi[0] - ox) * ix;
    const float t1y = (nd.lo[1] - oy) * iy, t2y = (nd.hi[1] - oy) * iy;
    const float t1z = (nd.lo[2] - oz) * iz, t2z = (nd.hi[2] - oz) * iz;
    const float te = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
    const float tx = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
    const float a = fmaxf(fmaxf(nd.lo[0] - h0, nd.lo[1] - h1), nd.lo[2] - h2);
    const float b = fmaxf(fmaxf(s0 - nd.hi[0], s1 - nd.hi[1]), s2 - nd.hi[2]);
    const float c = fmaxf(te, __int_as_float(1)) - tx;
    return fmaxf(fmaxf(a, b), c) <= 0.0f;
}

// Hit leaves of one walk are queued, up to kLeafQueue, and processed together in the
// order found.  A bounce's ray is fixed and the walk does not read the packet weight, so
// deferring a leaf changes no result; lanes holding a leaf keep stepping instead of
// idling until their neighbours find one.  The queue lives in LDS, one column per lane
// (slot q of lane t at [q][t]: 64 consecutive words per wave, no bank conflicts), so
// its depth costs no registers.  Measured on the Tri-X frame (transport ms): register
// queue 1961 / 1626 / 1514 / 1460 with depth 1 / 2 / 3 / 4 (30 VGPRs spilled at
// depth 4); LDS queue 1415 / 1378 / 1363 / 1366 / 1369 with depth 4 / 8 / 12 / 16 / 24
// (T-Max frame: 4027 / 3848 / 3809 / 3842 / 3865).  Every launch is 256 lanes per block.
// The leaf queue and the candidates of the absorption pre-test (below) are LDS columns,
// 1 KB per slot and 256-lane block.  Defaults (six waves per SIMD, the per-tile and
// replay kernels): 12 leaves and 8 candidates, 20 KB; the seven-wave frame transport: 10
// and 12, 22 KB, the most that seven blocks fit in a CU's 160 KB (Tri-X frame transport
// 996 ms with 12 and 8, 954 with 12 and 10, 946 with 10 and 12, 954 with 8 and 14; T-Max
// at six waves 2912, 2908, 2928 and 2974 ms, and 2911 with 12 and 12).
constexpr int kLeafQueue = 12;
constexpr uint32_t kCandidates = 8;
constexpr int kLeafQueueW7 = 10;
constexpr uint32_t kCandidatesW7 = 12;
constexpr uint32_t kQueueLanes = 256;

// kMode 0: latent counts (the transport); 1: expected-dose replay; 2: gradient replay
// (single tile only), with R the replay outputs and sums the lane's parameter sums.
enum : int { kCounts = 0, kReplayDose = 1, kReplayGrad = 2 };

// kPreRec (frame launches): the pre-test reads the crystals' 16-byte PreRec records pre
// (one load per crystal) instead of the centre and semi-axes of the 48-byte records.
template <bool kFrame, int kMode = kCounts, bool kPre = false, int kLQ = kLeafQueue, uint32_t kCand = kCandidates,
          bool kPreRec = false>
__device__ __forceinline__ void transport_persistent(const TransportArgs &C, const TileDev *tiles,
                                                     uint32_t n_tiles, const Node *nodes,
                                                     const float4 *recs, uint32_t *counts,
                                                     const float *emap, unsigned long long *work,
                                                     uint64_t total, LaneCounters &cnt,
                                                     const ReplayArgs *R = nullptr,
                                                     ReplaySums *sums = nullptr,
                                                     const PreRec *pre = nullptr)
{
    static_assert(kMode == kCounts || !kFrame, "replay runs on one tile");
    // the leaf cursor keeps the next queue slot (up to kLQ) in the top 4 bits of qk
    static_assert(kLQ >= 1 && kLQ <= 15, "leaf queue depth must fit the cursor's 4-bit slot field");
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t chunk = max(C.chunk, 64u);
    unsigned long long pool = 0;
    uint32_t pool_left = 0;
    // per wave: the tile of its latest chunk and that tile's packet range (LDS, not live
    // in registers between packet starts)
    __shared__ unsigned long long wrange[kQueueLanes / 64][2];
    __shared__ uint32_t wtile[kQueueLanes / 64];
    const uint32_t wv = threadIdx.x >> 6;
    if constexpr (kFrame) {
        wrange[wv][0] = 0ull;
        wrange[wv][1] = 0ull;
        wtile[wv] = 0u;
        __builtin_amdgcn_wave_barrier();
    }
    uint32_t phase = kNeed;
    const Node *nodes_b = kFrame ? nodes : C.nodes;
    const float4 *recs_b = kFrame ? recs : C.recs;
    uint32_t *counts_b = kFrame ? counts : C.counts;
    const __amdgpu_buffer_rsrc_t nodes_r = buffer_of(nodes_b), recs_r = buffer_of(recs_b), pre_r = buffer_of(pre);
    LaneTile T;
    PacketStart S0;
    if (!kFrame) lane_tile_single(C, T, S0);
    PhotonRng rng;
    float ox = 0.f, oy = 0.f, oz = 0.f, dx = 0.f, dy = 0.f, dz = 1.f, w = 0.f;
    float step = 0.f;
    uint32_t b = 0, i = 0, nq = 0;   // nq: leaves queued in this lane's column of lq
    __shared__ uint32_t lq_all[kLQ][kQueueLanes];
    uint32_t *lq = &lq_all[0][threadIdx.x];
    __shared__ uint32_t cq_all[kCand][kQueueLanes];
    uint32_t *cq = &cq_all[0][threadIdx.x];
    // leaf cursor of the lane's queue: kc the next crystal of the open leaf, qk = (next
    // queue slot) << 28 | (the open leaf's end); a tile's records are addressed by 32-bit
    // byte offsets, so its crystal indices (and leaf ends) stay below 2^32 / 48 < 2^28
    uint32_t kc = 0u, qk = 0u;
    // per-wave words of the deposit pre-reduction (wave_aggregated_add)
    __shared__ uint32_t agg_u32[kQueueLanes / 64];
    __shared__ double agg_f64[kQueueLanes / 64];
    if constexpr (kPre) {
        if ((threadIdx.x & 63) == 0) {
            agg_u32[threadIdx.x >> 6] = 0u;
            agg_f64[threadIdx.x >> 6] = 0.0;
        }
        __builtin_amdgcn_wave_barrier();
    }
    // replay: exposure of the packet's pixel, relative weight tangents, HG score and the
    // packet's gradient sums (flushed when the lane takes its next packet)
    float E = 0.f, w_mu = 0.f, w_r = 0.f, score = 0.f, a_e = 0.f, a_mu = 0.f, a_g = 0.f, a_r = 0.f;
    size_t e_idx = 0;
    bool active = false;
    const bool prune = C.reach >= 0.0f;
    for (;;) {
        if constexpr (kMode == kReplayGrad) {
            if (phase == kNeed && active) {
                if (a_e != 0.0f) atomicAdd(&R->gmap[e_idx], (double)a_e);
                sums->mu += a_mu;
                sums->g += a_g;
                sums->r += a_r;
                active = false;
            }
        }
        // ---- hand out packets to lanes without one
        const unsigned long long nm = __ballot(phase == kNeed);
        if (nm) {
            // the wave's packet ids come in consecutive chunks, far smaller than a tile: the
            // tile of the latest chunk's first id is kept wave-uniform (wt, [wbase, wend)),
            // and a lane whose id lies in it skips the binary search
            // n lanes need a packet; r = how many of them sit below this lane (mbcnt)
            const uint32_t n = (uint32_t)__popcll(nm);
            const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(nm >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)nm, 0u));
            unsigned long long id = pool + r;
            if (n > pool_left) {
                unsigned long long nb = 0;
                if (lane == 0) nb = atomicAdd(work, (unsigned long long)chunk);
                // the pool is wave-uniform: keep it in scalar registers (readfirstlane
                // returns int: each half goes through uint32_t, or a low half of 2^31 or
                // more would be sign-extended into the high half)
                const uint32_t nb_hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(nb >> 32));
                const uint32_t nb_lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)nb);
                nb = ((unsigned long long)nb_hi << 32) | (unsigned long long)nb_lo;
                if constexpr (kFrame) {
                    if (nb < total) {
                        const uint32_t wt = tile_of_packet(tiles, n_tiles, nb);
                        wrange[wv][0] = tiles[wt].photon_base;
                        wrange[wv][1] = wt + 1u < n_tiles ? tiles[wt + 1u].photon_base : total;
                        wtile[wv] = wt;
                        __builtin_amdgcn_wave_barrier();
                    }
                }
                if (r >= pool_left) id = nb + (r - pool_left);
                pool = nb + (n - pool_left);
                pool_left = chunk - (n - pool_left);
            } else {
                pool += n;
                pool_left -= n;
            }
            cnt.packets += (unsigned long long)__popcll(__builtin_amdgcn_ballot_w64(phase == kNeed && id < total));
            if (phase == kNeed) {
                if (id >= total) {
                    phase = kDone;
                } else {
                    uint64_t p = id;
                    PacketStart S = S0;
                    if constexpr (kFrame) {
                        const uint32_t ti = (id >= wrange[wv][0] && id < wrange[wv][1]) ? wtile[wv]
                                                                                     : tile_of_packet(tiles, n_tiles, id);
                        lane_tile_fill(tiles[ti], emap, C.estride, T, S);
                        p = id - tiles[ti].photon_base;
                    }
                    rng.init(p, S.key0, S.key1);
                    float x, y;
                    if (C.ppp > 0) {
                        // a tile's packet ids fit 32 bits (2^32 packets would be 3.4e7 pixels
                        // at 128 per pixel), where the divisions are a few instructions
                        // instead of a 64-bit software division each
                        uint32_t px, py;
                        if (p < (1ull << 32)) {
                            const uint32_t pix = (uint32_t)p / C.ppp;
                            px = pix % S.tw;
                            py = pix / S.tw;
                        } else {
                            const uint64_t pix = p / C.ppp;
                            px = (uint32_t)(pix % S.tw);
                            py = (uint32_t)(pix / S.tw);
                        }
                        const float ux = rng.u(), uy = rng.u();
                        x = ((float)px + ux) * S.pw;
                        y = ((float)py + uy) * S.ph;
                    } else {
                        const float ux = rng.u(), uy = rng.u();
                        x = ux * T.W;
                        y = uy * T.H;
                    }
                    w = 1.0f;
                    if (S.emap) {
                        const float uu = fminf(fmaxf(x / T.W, 0.0f), 1.0f);
                        const float vv = fminf(fmaxf(y / T.H, 0.0f), 1.0f);
                        const uint32_t ex = min((uint32_t)roundf(uu * (float)(S.tw - 1)), S.tw - 1);
                        const uint32_t ey = min((uint32_t)roundf(vv * (float)(S.th - 1)), S.th - 1);
                        e_idx = (size_t)ey * C.estride + ex;
                        w = S.emap[e_idx];
                    }
                    w = fmaxf(w, 0.0f);
                    E = w;
                    if (w > 0.0f) {
                        const float th = rng.u() * C.theta_max;
                        const float ph = rng.u() * kTwoPi;
                        const float st = __sinf(th);
                        dx = st * __cosf(ph);
                        dy = st * __sinf(ph);
                        dz = __cosf(th);
                        normalize3(dx, dy, dz);
                        const float cz = fminf(fmaxf(fabsf(dz), 0.0f), 1.0f);
                        const float m = 1.0f - cz;
                        const float fr = C.r0 + (1.0f - C.r0) * (m * m * m * m * m);
                        if (!(rng.u() < fr)) {
                            ox = x;
                            oy = y;
                            oz = 0.001f;
                            b = 0;
                            phase = kStart;
                            if constexpr (kMode != kCounts) {
                                w = 1.0f;
                                w_mu = w_r = score = 0.0f;
                                a_e = a_mu = a_g = a_r = 0.0f;
                                active = true;
                            }
                        }
                    }
                }
            }
        }
        if (__ballot(phase != kDone) == 0ull) break;
        // ---- start a bounce: free path, step, segment box
        bool started = false;
        if (phase == kStart) {
            if (b >= C.max_bounces) {
                phase = kNeed;
            } else {
                b++;
                started = true;
                float s;
                if (C.mu_t <= 0.0f) s = 1e6f;
                else s = -__logf(1.0f - rng.u()) / C.mu_t;
                step = fminf(s, C.step_cap);
                i = 0;
                nq = 0u;
                phase = kWalk;
            }
        }
        cnt.bounces += (unsigned long long)__popcll(__ballot(started));
        // ---- node steps (reference DFS order via skip links)
        {
            // wave-uniform counts in scalar registers, and a lane's step condition as one
            // compare: lim is the node count while the lane walks with room in its queue,
            // else 0 (a compound condition costs the ballot two more VALU operations)
            const uint32_t walking = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(phase == kWalk));
            const uint32_t stop =
                __builtin_amdgcn_readfirstlane((walking * (100u - min(C.wait_pct, 100u))) / 100u);
            uint32_t lim = phase == kWalk && nq == 0u ? T.n_nodes : 0u;
            // the first step runs whenever a lane can step, later ones while more than stop
            // lanes can (0 lanes is at most stop)
            uint32_t ns = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(i < lim));
            uint32_t steps = 0;
            // the walk's inverse direction and segment box, from the bounce's origin,
            // direction and step on every pass (not kept live through the leaf phase)
            float ix = 0.f, iy = 0.f, iz = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f, h0 = 0.f, h1 = 0.f, h2 = 0.f;
            if (ns != 0u) {
                ix = 1.0f / dx;
                iy = 1.0f / dy;
                iz = 1.0f / dz;
                const float ex = ox + dx * step, ey = oy + dy * step, ez = oz + dz * step;
                // without pruning the segment box is unbounded and every node is near
                const float mg = prune ? C.reach + 1e-3f + 1e-5f * (fabsf(ox) + fabsf(oy) + fabsf(oz) + step)
                                       : __builtin_inff();
                s0 = fminf(ox, ex) - mg;
                h0 = fmaxf(ox, ex) + mg;
                s1 = fminf(oy, ey) - mg;
                h1 = fmaxf(oy, ey) + mg;
                s2 = fminf(oz, ez) - mg;
                h2 = fmaxf(oz, ez) + mg;
            }
            while (ns != 0u) {
                steps += ns;
                if (i < lim) {
                    const NodeView nd = load_node_buf(nodes_r, T.node_base + i);
                    const bool hit = near_box_hit(nd, s0, s1, s2, h0, h1, h2, ox, oy, oz, ix, iy, iz);
                    // a stepping lane has room in its queue (lim drops to 0 when it fills),
                    // so the slot is written unconditionally and taken only by a hit leaf
                    lq[kQueueLanes * nq] = nd.prim;
                    nq += (hit & (nd.prim != kInterior)) ? 1u : 0u;
                    lim = nq == (uint32_t)kLQ ? 0u : lim;
                    // a leaf's skip link is its own index + 1 in the threaded layout, so a
                    // hit node of either kind continues at i + 1
                    i = hit ? i + 1 : nd.skip;
                }
                ns = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(i < lim));
                if (ns <= stop) break;
            }
            cnt.nodes += steps;
        }
        // ---- the queued leaves, in the order found: Beer-Lambert on each crystal hit
        // before the step end, in rounds of two phases.  A: the lane runs the cheap
        // conservative pre-test (crystal_may_absorb) over its leaves' crystals in order
        // from its cursor (the open leaf's crystals left, then the next queue slots),
        // keeping the candidates in an LDS column; a full column ends its phase A.
        // B: it tests the candidates in the same order with crystal_hit.  A crystal the
        // pre-test drops cannot record an event, and no draw or weight depends on the
        // dropped ones, so every result is that of testing all of them.  A lane whose
        // leaves are not all done keeps its queue and cursor, takes no node steps in the
        // next pass of the loop and continues there.
        uint32_t nc = 0u;
        bool more = phase == kWalk && nq > 0u;
        while (__builtin_amdgcn_ballot_w64(more && nc < kCand) != 0ull) {
            if (more && nc < kCand) {
                // open the next leaf with crystals, or end at the queue's end
                while (kc == (qk & 0x0FFFFFFFu) && more) {
                    const uint32_t qd = qk >> 28;
                    if (qd < nq) {
                        const uint32_t lf = lq[kQueueLanes * qd];
                        kc = lf >> 4;
                        qk = ((qd + 1u) << 28) | (kc + (lf & 15u));
                    } else {
                        more = false;
                    }
                }
                if (more) {
                    // two crystals at a time while the column has room for both (their loads
                    // in flight together): the open leaf's next two, or its last one and the
                    // first of the next queued leaf
                    bool two = nc + 1u < kCand;
                    uint32_t k2 = kc + 1u, qk2 = qk;
                    if (two && k2 == (qk & 0x0FFFFFFFu)) {
                        const uint32_t qd = qk >> 28;
                        two = qd < nq;
                        if (two) {
                            const uint32_t lf = lq[kQueueLanes * qd];
                            k2 = lf >> 4;
                            qk2 = ((qd + 1u) << 28) | (k2 + (lf & 15u));
                            two = (lf & 15u) != 0u;   // a leaf of no crystals: one this step
                        }
                    }
                    bool m1, m2;
                    if constexpr (kPreRec) {
                        const uint32_t o = (T.crystal_base + kc) * 16u;
                        const uint32_t o2 = two ? (T.crystal_base + k2) * 16u : o;
                        const float4 a = buf_load4(pre_r, o), a2 = buf_load4(pre_r, o2);
                        const PreRec r1 = {{a.x, a.y, a.z}, __float_as_uint(a.w)};
                        const PreRec r2 = {{a2.x, a2.y, a2.z}, __float_as_uint(a2.w)};
                        m1 = crystal_may_absorb_pre(r1, ox, oy, oz, dx, dy, dz, step);
                        m2 = two && crystal_may_absorb_pre(r2, ox, oy, oz, dx, dy, dz, step);
                    } else {
                        const uint32_t o = (T.crystal_base + kc) * 48u;
                        const uint32_t o2 = two ? (T.crystal_base + k2) * 48u : o;
                        const float4 ca = buf_load4(recs_r, o), ch = buf_load4(recs_r, o + 32u);
                        const float4 ca2 = buf_load4(recs_r, o2), ch2 = buf_load4(recs_r, o2 + 32u);
                        const float cc[3] = {ca.x, ca.y, ca.z}, hh[3] = {ch.x, ch.y, ch.z};
                        const float cc2[3] = {ca2.x, ca2.y, ca2.z}, hh2[3] = {ch2.x, ch2.y, ch2.z};
                        m1 = crystal_may_absorb(cc, hh, __float_as_uint(ch.w), ox, oy, oz, dx, dy, dz, step);
                        m2 = two && crystal_may_absorb(cc2, hh2, __float_as_uint(ch2.w), ox, oy, oz, dx, dy, dz, step);
                    }
                    cq[kQueueLanes * nc] = m1 ? kc : k2;
                    nc += m1 ? 1u : 0u;
                    if (m1 && m2) cq[kQueueLanes * nc] = k2;
                    nc += m2 ? 1u : 0u;
                    cnt.prims += two ? 2u : 1u;
                    if (two) {
                        kc = k2 + 1u;
                        qk = qk2;
                    } else {
                        kc++;
                    }
                }
            }
        }
        if (phase == kWalk && nq > 0u && !more) {
            nq = 0u;   // every queued leaf done: the walk goes on
            kc = qk = 0u;
        }
#pragma unroll 1
        for (uint32_t j = 0; j < kCand; j++) {
            if (j >= nc) continue;
            const uint32_t k = cq[kQueueLanes * j];
            const Rec r = load_rec_buf(recs_r, T.crystal_base + k);
            float t0, t1;
            if (!crystal_hit(r.c, r.q, r.h, r.morph, ox, oy, oz, dx, dy, dz, t0, t1)) continue;
            if (!(t0 < step)) continue;
            const float L = fmaxf(fminf(t1, step) - t0, 0.0f);
            if (!(L > 0.0f)) continue;
            const float pa = 1.0f - __expf(-C.mu_abs * L);
            if constexpr (kMode == kCounts) {
                const float lam = w * pa * r.sens * 5.0f;
                const float fl = fminf(floorf(lam), 4294967040.0f);
                const uint32_t whole = fl > 0.0f ? (uint32_t)fl : 0u;
                const float frac = fminf(fmaxf(lam - (float)whole, 0.0f), 1.0f);
                const uint32_t k_lat = whole + (rng.u() < frac ? 1u : 0u);
                if constexpr (kPre) {
                    cnt.atomics += wave_aggregated_add(counts_b, T.crystal_base + k, k_lat, k_lat > 0u,
                                                       &agg_u32[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)]);
                } else if (k_lat > 0u) {
                    atomicAdd(&counts_b[T.crystal_base + k], k_lat);
                }
                if (k_lat > 0u) {
                    cnt.abs++;
                    cnt.tot += k_lat;
                }
                w *= fmaxf(1.0f - pa * 0.9f, 0.05f);
            } else {
                cnt.abs++;
                if constexpr (kMode == kReplayGrad) {
                    const float base = 5.0f * r.sens * R->gdose[k];
                    const float dep = w * pa;
                    a_e += base * dep;
                    a_mu += base * E * (w_mu * pa + w * L * (1.0f - pa));
                    a_r += base * E * (w_r * pa);
                    a_g += base * E * dep * score;
                } else {
                    const double dep = (double)(((E * w) * pa) * r.sens * 5.0f);
                    if constexpr (kPre) {
                        cnt.atomics += wave_aggregated_add(R->dose, k, dep, true,
                                                           &agg_f64[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)]);
                    } else {
                        atomicAdd(&R->dose[k], dep);
                    }
                }
                const float m = 1.0f - pa * 0.9f;
                const float dm = -0.9f * L * (1.0f - pa);
                w_mu = w_mu * m + w * dm;
                w_r = w_r * m;
                w = w * m;
            }
        }
        // ---- end of the bounce: move, boundaries, halation or scattering, roulette
        if (phase == kWalk && nq == 0u && i >= T.n_nodes) {
            const float qx = ox + dx * step, qy = oy + dy * step, qz = oz + dz * step;
            if (qx < 0.0f || qx >= T.W || qy < 0.0f || qy >= T.H || qz <= 0.0f) {
                phase = kNeed;
            } else if (qz >= C.D) {
                ox = qx;
                oy = qy;
                oz = C.D - 0.001f;
                const float inx = dx, iny = dy, inz = dz;
                phase = kStart;
                if constexpr (kMode == kCounts) {
                    w *= C.hal_eff;
                    if (w < 1e-4f) {
                        if (rng.u() < 0.1f) w /= 0.1f;
                        else phase = kNeed;
                    }
                } else {
                    w_r = w_r * C.hal_eff + w * R->hal_unit;
                    w_mu = w_mu * C.hal_eff;
                    w = w * C.hal_eff;
                }
                if (phase == kStart) {
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
                    if constexpr (kMode == kCounts) {
                        if (!(w > 0.0f)) phase = kNeed;
                    } else {
                        if (w < R->cut && w_r < R->cut) phase = kNeed;
                    }
                    if (phase == kStart && C.base > 0.0f) {
                        if (!(inz > 1e-6f) || !(-dz > 1e-6f)) {
                            phase = kNeed;
                        } else {
                            ox += C.base * (inx / inz + dx / (-dz));
                            oy += C.base * (iny / inz + dy / (-dz));
                            if (ox < 0.0f || ox >= T.W || oy < 0.0f || oy >= T.H) phase = kNeed;
                        }
                    }
                }
            } else {
                ox = qx;
                oy = qy;
                oz = qz;
                const float xi1 = rng.u(), xi2 = rng.u();
                float ct;
                if (fabsf(C.g) < 1e-4f) {
                    ct = 2.0f * xi1 - 1.0f;
                } else {
                    const float term = (1.0f - C.g * C.g) / (1.0f - C.g + 2.0f * C.g * xi1);
                    ct = (1.0f + C.g * C.g - term * term) / (2.0f * C.g);
                }
                ct = fminf(fmaxf(ct, -1.0f), 1.0f);
                if constexpr (kMode == kReplayGrad) score += hg_score(C.g, ct);
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
                w *= C.albedo;
                phase = kStart;
                if constexpr (kMode == kCounts) {
                    if (w < 1e-3f) {
                        if (rng.u() < 0.1f) w /= 0.1f;
                        else phase = kNeed;
                    }
                } else {
                    w_mu *= C.albedo;
                    w_r *= C.albedo;
                    if (w < R->cut && w_r < R->cut) phase = kNeed;
                }
            }
        }
    }
    // without the pre-reduction every absorbing hit (every event of a dose replay) is one atomic
    if constexpr (!kPre) cnt.atomics = cnt.abs;
}

// Path length through developed silver of the ray at (ox, oy) along -z: for a ray
// along -z the reference's slab test (+inf inverse components, NaN-ignoring min/max)
// is exactly "strictly inside the node's x-y rectangle", every box reaching below the
// origin; with outward-rounded float boxes and float origins the comparison is
// exact, so the same crystals are tested as in the double-precision reference.
// The walk is split into node steps and queued leaves, as in the transport: a lane
// queues the leaves it may hit (up to kScanQueue node indices, in an LDS column of its
// own, lq[q * kQueueLanes]) and keeps stepping; steps continue until 90 % of the walking
// lanes have a full queue or have ended their walk, then every lane tests the crystals of
// its queued leaves in the order found.  The steps read the 16-byte grid nodes (one load
// per step, ScanQNode), which enter every node the float test enters; a queued leaf is
// confirmed before its crystals by the grid's narrow test or, in the grid's rounding
// band, by its float box, so each ray meets the leaves of the float walk in its order.
// consts holds two float4 per record: its
// down_constants (the ray direction in the crystal frame is a constant of the crystal
// for rays along -z) and its centre's rounding residual, with which crystal_chord_down
// solves each crystal from the level of its centre.  The
// chords are added in depth-first order either way, so queueing changes no bit (frame scans 261 -> 148 ms Tri-X, 417 -> 191 ms
// T-Max against testing each leaf when it is hit; queue depths 8 / 12 / 16: 91.8 / 91.3 /
// 90.3 ms Tri-X, 113.8 / 112.2 / 111.8 ms T-Max).  Lanes with active = false take part
// in no step.
constexpr uint32_t kScanQueue = 16;

__device__ __forceinline__ float silver_path_queued(const Node *nodes, const ScanQNode *qnodes, uint32_t n_nodes,
                                                    const float4 *recs, const float4 *consts, float ox, float oy,
                                                    float oz, float bx, float by, float sc, bool active,
                                                    uint32_t *lq)
{
    float P = 0.0f;
    uint32_t i = active ? 0u : n_nodes, nq = 0;
    const __amdgpu_buffer_rsrc_t nodes_r = buffer_of(nodes), q_r = buffer_of(qnodes), recs_r = buffer_of(recs),
                                 consts_r = buffer_of(consts);
    const ScanQRay qr = scan_qray(ox, oy, bx, by, sc);   // the ray on the tree's grid (ScanQNode)
    for (;;) {
        const uint32_t walking = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(i < n_nodes));
        if (walking == 0u) break;
        const uint32_t stop = walking / 10u;
        uint32_t lim = n_nodes;   // 0 once the lane's queue is full
        for (uint32_t it = 0;; it++) {
            const bool stepping = i < lim;
            const uint32_t ns = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(stepping));
            if (ns == 0u || (it > 0u && ns <= stop)) break;
            if (stepping) {
                const BufWords q = __builtin_amdgcn_raw_buffer_load_b128(q_r, (int)(i << 4), 0, 0);
                const bool leaf = (uint32_t)q.z == kScanLeaf;
                const bool hit = scan_q_may((uint32_t)q.x, (uint32_t)q.y, qr);
                // a stepping lane has room in its queue, so the slot is written always and
                // taken only by a leaf
                lq[kQueueLanes * nq] = i;
                nq += (hit && leaf) ? 1u : 0u;
                lim = nq == kScanQueue ? 0u : lim;
                // a leaf's skip link is its index + 1
                i = (hit || leaf) ? i + 1 : (uint32_t)q.z;
            }
        }
#pragma unroll 1
        for (uint32_t q = 0; q < kScanQueue; q++) {
            if (q >= nq) continue;
            const uint32_t j = lq[kQueueLanes * q];   // the leaf's node
            const BufWords qn = __builtin_amdgcn_raw_buffer_load_b128(q_r, (int)(j << 4), 0, 0);
            bool in = scan_q_surely((uint32_t)qn.x, (uint32_t)qn.y, qr);
            if (!in) {
                const NodeView nd = load_node_buf(nodes_r, j);
                in = nd.lo[0] < ox && ox < nd.hi[0] && nd.lo[1] < oy && oy < nd.hi[1];
            }
            if (!in) continue;
            const uint32_t lf = (uint32_t)qn.w;
            const uint32_t first = lf >> 4, cnt = lf & 15u;
            for (uint32_t k = first; k < first + cnt; k++) {
                // a scan record's fourth word holds its bounding-sphere radius
                const float4 c0 = buf_load4(recs_r, k * 48u);
                if (!sphere_may_hit_z_r(c0, c0.w, ox, oy)) continue;
                const float4 kc = buf_load4(consts_r, k << 5), cl = buf_load4(consts_r, (k << 5) + 16u);
                const float4 c1 = buf_load4(recs_r, k * 48u + 16u), c2 = buf_load4(recs_r, k * 48u + 32u);
                const float cc[3] = {c0.x, c0.y, c0.z}, qq[4] = {c1.x, c1.y, c1.z, c1.w},
                            hh[3] = {c2.x, c2.y, c2.z};
                const float kk[4] = {kc.x, kc.y, kc.z, kc.w}, lo[3] = {cl.x, cl.y, cl.z};
                P += crystal_chord_down(cc, qq, hh, __float_as_uint(c2.w), kk, lo, ox, oy, oz);
            }
        }
        nq = 0u;
    }
    return P;
}

// Origin of ray (ax, ay, sx, sy) of pixel (px, py): stratified jitter of
// transmittance.rs:110-120 inside an ss x ss aperture.
__device__ __forceinline__ void scan_origin(uint32_t pix, uint32_t px, uint32_t py, uint32_t ax,
                                            uint32_t ay, uint32_t sx, uint32_t sy, uint32_t g,
                                            uint32_t ss, float pw, float ph, uint32_t key0,
                                            uint32_t key1, float &ox, float &oy)
{
    const uint32_t ray = ((ay * ss + ax) * g + sy) * g + sx;
    uint32_t c[4] = {pix, ray >> 1, kTagScan, 0u};
    philox(c, key0, key1);
    const uint32_t wx = (ray & 1u) ? c[2] : c[0];
    const uint32_t wy = (ray & 1u) ? c[3] : c[1];
    const float subx = ((float)sx + u24(wx)) / (float)g;
    const float suby = ((float)sy + u24(wy)) / (float)g;
    const float apx = ((float)ax + subx) / (float)ss;
    const float apy = ((float)ay + suby) / (float)ss;
    ox = ((float)px + apx) * pw;
    oy = ((float)py + apy) * ph;
}

// Lanes per pixel of the scan kernels: each lane takes rays_per_lane of the pixel's rays
// (1: one lane per ray, the default), the pixel's lanes adjacent in the wave; a shape
// that does not tile the wavefront falls back to one lane per ray when the rays divide
// the wavefront, else one lane per pixel.
FG_HD uint32_t scan_lanes(uint32_t rays, uint32_t rays_per_lane = 1u)
{
    const uint32_t r = rays_per_lane ? rays_per_lane : 1u;
    if (rays % r == 0u) {
        const uint32_t L = rays / r;
        if (L <= 64u && (64u % L) == 0u) return L;
    }
    return (rays <= 64u && (64u % rays) == 0u) ? rays : 1u;
}

// Mean transmittance of pixel (px, py) computed by L cooperating lanes (sub = lane index
// within the pixel's group); the group sum is a fixed xor-tree, so results are
// reproducible.  Every lane of the wave must call this (invalid lanes pass valid=false);
// blocks are kQueueLanes lanes.
__device__ __forceinline__ float scan_pixel(bool valid, uint32_t pix, uint32_t px, uint32_t py,
                                            uint32_t sub, uint32_t L, uint32_t g, uint32_t ss,
                                            float pw, float ph, float z0, float sigma, uint32_t key0,
                                            uint32_t key1, const Node *nodes, const ScanQNode *qnodes,
                                            uint32_t n_nodes, const float4 *recs, const float4 *consts)
{
    __shared__ uint32_t lq_all[kScanQueue][kQueueLanes];
    uint32_t *lq = &lq_all[0][threadIdx.x];
    const uint32_t R = g * g * ss * ss;
    float acc = 0.0f;
    if (valid) {
        float bx = 0.0f, by = 0.0f, sc = 1.0f;   // no node: every walk ends at once
        if (n_nodes > 0u) {
            const Node root = nodes[0];
            scan_qframe(root, bx, by, sc);
        }
        for (uint32_t r = sub; r < R; r += L) {
            const uint32_t sx = r % g, sy = (r / g) % g, ax = (r / (g * g)) % ss, ay = r / (g * g * ss);
            float ox, oy;
            scan_origin(pix, px, py, ax, ay, sx, sy, g, ss, pw, ph, key0, key1, ox, oy);
            acc += expf(-sigma * silver_path_queued(nodes, qnodes, n_nodes, recs, consts, ox, oy, z0, bx, by, sc,
                                                    true, lq));
        }
    }
    for (uint32_t off = L >> 1; off >= 1u; off >>= 1) acc += __shfl_xor(acc, (int)off, 64);
    return acc / (float)R;
}

// ScanMode mapping (transmittance.rs:144-152) and D_neg (image_pipeline.rs:1006-1011)
__device__ __forceinline__ float scan_value(float T, uint32_t mode)
{
    if (mode == 0u) return T;
    if (mode == 1u) return 1.0f - T;
    if (mode == 2u) {
        const float t = fmaxf(T, 1e-4f);
        return fminf(fmaxf(-log10f(t) / 3.0f, 0.0f), 1.0f);
    }
    const double t = fmin(fmax((double)T, 1e-4), 1.0);
    return (float)(-log10(t) + 0.05);
}

}  // namespace dev
}  // namespace fg
