// Device layouts and shared inline math of the MI355X photon-transport + scan path.
//
// Everything here is compiled both for the host (C++) and for gfx950 (HIP).  The
// floating-point operation order of the geometry functions is fixed (the library is
// built with -ffp-contract=off), so a host evaluation of the same function returns
// the same bits as the kernel.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FG_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define FG_HD inline
#endif

namespace fg {

// Philox counter layouts (key = (lo32(seed), hi32(seed))):
//   photon stream: ctr = (lo32(photon), hi32(photon), draw >> 2, kTagPhoton), word draw & 3
//   scan jitter:   ctr = (pixel, ray >> 1, kTagScan, 0), words 2*(ray & 1) (x) and +1 (y)
//   sequential:    ctr = (lo32(block), hi32(block), 0, kTagSeq); u64 = w0 | w1 << 32
//   dart:          ctr = (cell, round, kTagDart, 0), words 0..2 the candidate's jitter
//   crystal:       ctr = (lo32(point), hi32(point), block, kTagAttr), sequential words
constexpr uint32_t kTagPhoton = 0x50484f54u;
constexpr uint32_t kTagScan = 0x5343414eu;
constexpr uint32_t kTagSeq = 0x53455120u;
constexpr uint32_t kTagDart = 0x44415254u;
constexpr uint32_t kTagAttr = 0x41545452u;

// 32-byte BVH node in threaded ("skip-link") pre-order: a node's right subtree follows
// it directly and its left subtree follows the right one, so walking "hit -> i + 1,
// miss or leaf done -> skip" visits nodes in exactly the order of a depth-first
// traversal that pops the right child first, with no stack.  A leaf's subtree is the
// leaf alone, so its skip is its own index + 1 and the walk reduces to "hit -> i + 1,
// miss -> skip" for both kinds of node (both builders, host and device, keep this).
struct alignas(32) Node {
    float lo[3];
    float hi[3];
    uint32_t skip;   // next node when this subtree is finished or missed (n_nodes = end)
    uint32_t prim;   // kInterior, or (first_prim << 4) | count for a leaf
};
constexpr uint32_t kInterior = 0xFFFFFFFFu;
constexpr uint32_t kMaxLeafCount = 15u;

// 48-byte crystal record in BVH leaf order: three 16-byte vectors.
struct alignas(16) Crystal {
    float c[3];
    float sens;
    float q[4];      // unit quaternion (w, i, j, k)
    float h[3];      // semi-axes times filament expansion
    uint32_t morph;  // 0 cubic, 1 tabular, 2 octahedral
};

static_assert(sizeof(Node) == 32, "node must be 32 bytes");
static_assert(sizeof(Crystal) == 48, "crystal record must be 48 bytes");

// 16-byte node of the scan's walk along -z, one per Node at the same index: the node's
// x-y box on a 16-bit grid of the tree (scan_qframe), rounded outward, and the skip link
// of an interior node or kScanLeaf with the leaf's prim.  With X = (o - origin) * scale
// for a ray through (ox, oy), a point strictly inside the node's float box has
// lo < X < hi on both axes; the float evaluation x of X errs by less than 0.01 grid
// units for every ray of a tile (relative error 2^-23, |X| far below 10^5), so
// x >= lo - 0.5 and x <= hi + 0.5 hold for every such point (the walk may enter a few
// more interior nodes, never fewer), and x >= lo + 1.5, x <= hi - 1.5 (lo and hi
// within one unit of the exact box) prove the point strictly inside a leaf; in the band
// between, the walk tests the leaf's float box itself.  The ray's bounds are rounded to
// integers once (ScanQRay), so a node costs four integer comparisons.
struct alignas(16) ScanQNode {
    uint32_t lo;     // x | y << 16, rounded down
    uint32_t hi;     // x | y << 16, rounded up
    uint32_t word;   // skip link of an interior node, kScanLeaf for a leaf
    uint32_t prim;   // the node's prim
};
static_assert(sizeof(ScanQNode) == 16, "scan node must be 16 bytes");
constexpr uint32_t kScanLeaf = 0xFFFFFFFFu;

// grid of a tree: origin at the root box's low x-y corner, 65534 units across its wider
// x-y extent
FG_HD void scan_qframe(const Node &root, float &bx, float &by, float &s)
{
    bx = root.lo[0];
    by = root.lo[1];
    const float ext = fmaxf(root.hi[0] - root.lo[0], root.hi[1] - root.lo[1]);
    s = ext > 0.0f ? 65534.0f / ext : 1.0f;
}

// (v - b) * s is exact in double (float operands), so the rounding is exact too
FG_HD uint32_t scan_qcoord(float v, float b, float s, bool up)
{
    const double x = ((double)v - (double)b) * (double)s;
    const double r = up ? ceil(x) : floor(x);
    return (uint32_t)fmin(fmax(r, 0.0), 65535.0);
}

// A ray on the grid: integer bounds for the two node tests.  For an integer l,
// l <= floor(x + 0.5) exactly when x >= l - 0.5, and l < floor(x - 0.5) exactly when
// x >= l + 1.5 (likewise ceil for the upper bounds); the float sums round by less than
// 0.01 units.
struct ScanQRay {
    int32_t may_lo[2], may_hi[2];   // node enters when lo <= may_lo and hi >= may_hi
    int32_t in_lo[2], in_hi[2];     // leaf surely holds the point when lo < in_lo and hi > in_hi
};

FG_HD int32_t scan_qint(float v) { return (int32_t)fminf(fmaxf(v, -1e9f), 1e9f); }

FG_HD ScanQRay scan_qray(float ox, float oy, float bx, float by, float s)
{
    const float q[2] = {(ox - bx) * s, (oy - by) * s};
    ScanQRay r;
    for (int a = 0; a < 2; a++) {
        r.may_lo[a] = scan_qint(floorf(q[a] + 0.5f));
        r.may_hi[a] = scan_qint(ceilf(q[a] - 0.5f));
        r.in_lo[a] = scan_qint(floorf(q[a] - 0.5f));
        r.in_hi[a] = scan_qint(ceilf(q[a] + 0.5f));
    }
    return r;
}

// every point strictly inside the node's float box passes
FG_HD bool scan_q_may(uint32_t lo, uint32_t hi, const ScanQRay &r)
{
    return (int32_t)(lo & 0xFFFFu) <= r.may_lo[0] && (int32_t)(hi & 0xFFFFu) >= r.may_hi[0] &&
           (int32_t)(lo >> 16) <= r.may_lo[1] && (int32_t)(hi >> 16) >= r.may_hi[1];
}

// only points strictly inside the node's float box pass
FG_HD bool scan_q_surely(uint32_t lo, uint32_t hi, const ScanQRay &r)
{
    return (int32_t)(lo & 0xFFFFu) < r.in_lo[0] && (int32_t)(hi & 0xFFFFu) > r.in_hi[0] &&
           (int32_t)(lo >> 16) < r.in_lo[1] && (int32_t)(hi >> 16) > r.in_hi[1];
}

FG_HD ScanQNode scan_qnode(const Node &nd, float bx, float by, float s)
{
    ScanQNode o;
    o.lo = scan_qcoord(nd.lo[0], bx, s, false) | scan_qcoord(nd.lo[1], by, s, false) << 16;
    o.hi = scan_qcoord(nd.hi[0], bx, s, true) | scan_qcoord(nd.hi[1], by, s, true) << 16;
    o.word = nd.prim == kInterior ? nd.skip : kScanLeaf;
    o.prim = nd.prim;
    return o;
}

// ------------------------------------------------------------------ Philox4x32-10
FG_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

// Philox4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, SC'11).
FG_HD void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int r = 0; r < 10; r++) {
#if defined(__HIP_DEVICE_COMPILE__)
        asm volatile("" : "+v"(k0), "+v"(k1));
#endif
        const uint32_t h0 = mulhi32(0xD2511F53u, c[0]), l0 = 0xD2511F53u * c[0];
        const uint32_t h1 = mulhi32(0xCD9E8D57u, c[2]), l1 = 0xCD9E8D57u * c[2];
        const uint32_t n0 = h1 ^ c[1] ^ k0, n2 = h0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = l1;
        c[2] = n2;
        c[3] = l0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// 24-bit uniform in [0, 1), exactly representable in float
FG_HD float u24(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }

// ------------------------------------------------------------------ geometry (f32)
// UnitQuaternion::inverse_transform_vector (crystal.rs:151-154): rotation by the
// conjugate, t = 2 (u' x v), out = (t w + u' x t) + v with u' = -u.
FG_HD void qinv_rotate(const float q[4], float vx, float vy, float vz, float &ox, float &oy,
                       float &oz)
{
    const float w = q[0], ux = -q[1], uy = -q[2], uz = -q[3];
    const float tx = (uy * vz - uz * vy) * 2.0f;
    const float ty = (uz * vx - ux * vz) * 2.0f;
    const float tz = (ux * vy - uy * vx) * 2.0f;
    const float cx = uy * tz - uz * ty;
    const float cy = uz * tx - ux * tz;
    const float cz = ux * ty - uy * tx;
    ox = (tx * w + cx) + vx;
    oy = (ty * w + cy) + vy;
    oz = (tz * w + cz) + vz;
}

// HalideCrystal::intersect_ray (crystal.rs:140-217) on a crystal record; returns true
// and (t0, t1) = (max(t_enter, 0), t_exit) on a hit.
FG_HD bool crystal_hit(const float c[3], const float q[4], const float h[3], uint32_t morph,
                       float ox, float oy, float oz, float dx, float dy, float dz, float &t0,
                       float &t1)
{
    float lx, ly, lz, ex, ey, ez;
    qinv_rotate(q, ox - c[0], oy - c[1], oz - c[2], lx, ly, lz);
    qinv_rotate(q, dx, dy, dz, ex, ey, ez);
    if (morph == 0u) {
        const float ix = 1.0f / ex, iy = 1.0f / ey, iz = 1.0f / ez;
        const float a1 = (-h[0] - lx) * ix, a2 = (h[0] - lx) * ix;
        const float a3 = (-h[1] - ly) * iy, a4 = (h[1] - ly) * iy;
        const float a5 = (-h[2] - lz) * iz, a6 = (h[2] - lz) * iz;
        const float tmin = fmaxf(fmaxf(fminf(a1, a2), fminf(a3, a4)), fminf(a5, a6));
        const float tmax = fminf(fminf(fmaxf(a1, a2), fmaxf(a3, a4)), fmaxf(a5, a6));
        if (tmax >= tmin && tmax > 0.0f) {
            t0 = fmaxf(tmin, 0.0f);
            t1 = tmax;
            return true;
        }
        return false;
    }
    const float px = lx / h[0], py = ly / h[1], pz = lz / h[2];
    const float qx = ex / h[0], qy = ey / h[1], qz = ez / h[2];
    const float a = qx * qx + qy * qy + qz * qz;
    const float b = 2.0f * (px * qx + py * qy + pz * qz);
    const float cc = (px * px + py * py + pz * pz) - 1.0f;
    const float disc = b * b - 4.0f * a * cc;
    if (disc < 0.0f) return false;
    const float s = sqrtf(disc);
    const float r0 = (-b - s) / (2.0f * a);
    const float r1 = (-b + s) / (2.0f * a);
    const float te = fminf(r0, r1), tx = fmaxf(r0, r1);
    if (tx > 0.0f) {
        t0 = fmaxf(te, 0.0f);
        t1 = tx;
        return true;
    }
    return false;
}

// Reciprocal for error margins: one hardware instruction on the device (1 ulp), the
// division on the host; either is far inside the margins it scales.
FG_HD float margin_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// Conservative pre-test of an absorption event on the segment o + t d, t in [0, step]
// (|d| = 1): false only when crystal_hit, evaluated in float on the same inputs, cannot
// return a hit with t0 < step and a chord L = min(t1, step) - t0 > 0.  The crystal lies
// in its bounding sphere (radius sqrt(3) hmax for the cube, hmax for the ellipsoid), and
// the test rejects a segment that misses that sphere by more than crystal_hit's float
// error can move its answer:
//  * cube: the slab times are quotients of differences that carry a few ulps of
//    |o - c| + R, so the hit decision and t0, t1 are those of a geometry perturbed by a
//    few ulps of that size; margin 1e-5 (|o - c|_1 + R), about 100 times that;
//  * ellipsoid: the discriminant b^2 - 4 a (|p|^2 - 1) in semi-axis units (p the origin,
//    |p| <= rho = (|o - c|_1 + R) / hmin) loses about 10 ulps of b^2 = 4 |p|^2 a to
//    cancellation, so d_perp^2 - 1 (d_perp: distance of the line from the centre in those
//    units) is decided within 6e-7 rho^2; the test keeps every segment with physical
//    distance^2 <= R^2 (1 + 1e-5 rho + 6e-6 rho^2), ten times that (and 1e-5 rho for the
//    rotation and division errors).  Near tangency the roots spread by at most
//    sqrt(error) / 2a <= 7.7e-4 rho hmax along the ray; the ends of the segment get 8e-3
//    rho hmax.
// Comparisons are written so that a NaN keeps the crystal.  Every quantity grows with
// hmax and with 1 / hmin, so the test keeps at least the same segments when it is given a
// larger hmax or a smaller hmin (crystal_may_absorb_pre).
FG_HD bool crystal_may_absorb_hm(const float c[3], float hmax, float hmin, bool cube, float ox, float oy, float oz,
                                 float dx, float dy, float dz, float step)
{
    const float px = c[0] - ox, py = c[1] - oy, pz = c[2] - oz;
    const float pn = fabsf(px) + fabsf(py) + fabsf(pz);
    float R, r2, mt;
    if (cube) {
        R = hmax * 1.7320525f;   // sqrt(3) (1 + 1e-6)
        const float m = 1e-5f * (pn + R);
        r2 = (R + m) * (R + m);
        mt = m;
    } else {
        R = hmax * 1.000001f;
        const float rho = (pn + R) * margin_rcp(hmin);
        r2 = (R * R) * (1.0f + rho * (1e-5f + 6e-6f * rho));
        mt = hmax * (8e-3f * rho) + 1e-5f * (pn + R);
    }
    const float ax = py * dz - pz * dy, ay = pz * dx - px * dz, az = px * dy - py * dx;
    const float tp = px * dx + py * dy + pz * dz;
    if (ax * ax + ay * ay + az * az > r2) return false;   // the line passes outside
    if (tp + (R + mt) <= 0.0f) return false;               // behind the origin
    if (tp - (R + mt) >= step) return false;               // beyond the segment's end
    return true;
}

FG_HD bool crystal_may_absorb(const float c[3], const float h[3], uint32_t morph, float ox, float oy, float oz,
                              float dx, float dy, float dz, float step)
{
    return crystal_may_absorb_hm(c, fmaxf(fmaxf(h[0], h[1]), h[2]), fminf(fminf(h[0], h[1]), h[2]), morph == 0u, ox,
                                 oy, oz, dx, dy, dz, step);
}

// binary16 bits of a float v >= 0, rounded up or down; only normal values, 0 and infinity
// result (below 2^-14 v rounds down to 0 and up to 2^-14, above 65504 down to 65504 and up
// to infinity)
FG_HD uint32_t f16_bits_dir(float v, bool up)
{
    if (!(v > 0.0f)) return 0u;
    if (v < 0x1p-14f) return up ? (1u << 10) : 0u;
    if (v > 65504.0f) return up ? 0x7C00u : 0x7BFFu;
    const uint32_t b = __builtin_bit_cast(uint32_t, v);
    const uint32_t h = ((((b >> 23) & 255u) - 112u) << 10) | ((b >> 13) & 1023u);
    return h + ((up && (b & 8191u)) ? 1u : 0u);   // a carry moves to the next binade exactly
}

// the value of binary16 bits made by f16_bits_dir (exact; on the device one conversion
// instruction, which is exact for every binary16 value)
FG_HD float f16_bits_value(uint32_t h)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
#endif
    const uint32_t e = (h >> 10) & 31u, m = h & 1023u;
    const uint32_t b = e == 0u ? 0u : (e == 31u ? 0x7F800000u : ((e + 112u) << 23) | (m << 13));
    return __builtin_bit_cast(float, b);
}

// 16-byte record of the frame transport's absorption pre-test, one per crystal at the
// record's index: the centre, the largest semi-axis rounded up and the smallest rounded
// down to binary16, and whether the crystal takes the ellipsoid branch.  The pre-test on
// it keeps every segment crystal_may_absorb keeps (larger hmax, smaller hmin), in one
// load instead of two.
struct alignas(16) PreRec {
    float c[3];
    uint32_t hw;   // up(hmax) | down(hmin) << 16 | (morph != 0) << 31
};
static_assert(sizeof(PreRec) == 16, "pre-test record must be 16 bytes");

FG_HD PreRec pre_record(const float c[3], const float h[3], uint32_t morph)
{
    PreRec r;
    r.c[0] = c[0];
    r.c[1] = c[1];
    r.c[2] = c[2];
    r.hw = f16_bits_dir(fmaxf(fmaxf(h[0], h[1]), h[2]), true) |
           f16_bits_dir(fminf(fminf(h[0], h[1]), h[2]), false) << 16 | (morph != 0u ? 1u : 0u) << 31;
    return r;
}

FG_HD bool crystal_may_absorb_pre(const PreRec &r, float ox, float oy, float oz, float dx, float dy, float dz,
                                  float step)
{
    return crystal_may_absorb_hm(r.c, f16_bits_value(r.hw & 0xFFFFu), f16_bits_value((r.hw >> 16) & 0x7FFFu),
                                 (r.hw >> 31) == 0u, ox, oy, oz, dx, dy, dz, step);
}

// Constants of a crystal for rays along -z (the scan): the direction (0, 0, -1) rotated
// into the crystal frame depends on the crystal only, so its inverse (cube) or its
// components over the semi-axes (other shapes) are computed once per crystal; k[3] is the
// bounding-sphere radius of the scan's pre-test.
FG_HD void down_constants(const float q[4], const float h[3], uint32_t morph, float k[4])
{
    float ex, ey, ez;
    qinv_rotate(q, 0.0f, 0.0f, -1.0f, ex, ey, ez);
    if (morph == 0u) {
        k[0] = 1.0f / ex;
        k[1] = 1.0f / ey;
        k[2] = 1.0f / ez;
    } else {
        k[0] = ex / h[0];
        k[1] = ey / h[1];
        k[2] = ez / h[2];
    }
    const float hm = fmaxf(fmaxf(h[0], h[1]), h[2]);
    const float r2 = morph == 0u ? h[0] * h[0] + h[1] * h[1] + h[2] * h[2] : hm * hm;
    k[3] = sqrtf(r2);
}

// Chord max(t_exit - max(t_enter, 0), 0) of a ray (ox, oy, oz) + t (0, 0, -1) through a
// crystal (HalideCrystal::intersect_ray, crystal.rs:140-217, as transmittance.rs:
// 141-150 uses it), 0 on a miss; k = down_constants, lo = the centre's rounding residual
// (f64 centre minus the record's f32 centre), so the ray's offset from the centre is
// (o - c) - lo, near the f64 offset (a record's f32 centre alone is off by up to half
// an ulp of its coordinate, 4e-6 at 40 units, which moves a ray grazing a rounded
// crystal across its rim).  The ray is solved from the point level with the crystal's
// centre, t = t' + (oz - c_z): scan rays start above the slab, tens of semi-axes away,
// and solving from there in f32 loses the chord of a thin crystal to cancellation (the
// ellipsoid's b^2 - 4ac cancels to 1e-3 of the chord for an 8:1 platelet 20 units below
// the origin).  From the centre's level the ellipsoid is solved about the ray's closest
// approach in the unit frame, m = p + t_c q with t_c = -(p.q) / (q.q), chord
// 2 sqrt((1 - m.m) / (q.q)); the box by its slabs.
FG_HD float crystal_chord_down(const float c[3], const float q[4], const float h[3], uint32_t morph,
                               const float k[4], const float lo[3], float ox, float oy, float oz)
{
    const float shift = (oz - c[2]) - lo[2];
    float lx, ly, lz;
    qinv_rotate(q, (ox - c[0]) - lo[0], (oy - c[1]) - lo[1], 0.0f, lx, ly, lz);
    float te, tx;
    if (morph == 0u) {
        const float ix = k[0], iy = k[1], iz = k[2];
        const float a1 = (-h[0] - lx) * ix, a2 = (h[0] - lx) * ix;
        const float a3 = (-h[1] - ly) * iy, a4 = (h[1] - ly) * iy;
        const float a5 = (-h[2] - lz) * iz, a6 = (h[2] - lz) * iz;
        te = fmaxf(fmaxf(fminf(a1, a2), fminf(a3, a4)), fminf(a5, a6));
        tx = fminf(fminf(fmaxf(a1, a2), fmaxf(a3, a4)), fmaxf(a5, a6));
        if (!(tx >= te)) return 0.0f;
    } else {
        const float px = lx / h[0], py = ly / h[1], pz = lz / h[2];
        const float qx = k[0], qy = k[1], qz = k[2];
        const float qq = qx * qx + qy * qy + qz * qz;
        const float tc = -(px * qx + py * qy + pz * qz) / qq;
        const float mx = px + tc * qx, my = py + tc * qy, mz = pz + tc * qz;
        const float rem = 1.0f - (mx * mx + my * my + mz * mz);
        if (!(rem >= 0.0f)) return 0.0f;
        const float half = sqrtf(rem / qq);
        te = tc - half;
        tx = tc + half;
    }
    // t = t' + shift: exit beyond the origin, entry clamped at the origin
    if (!(tx > -shift)) return 0.0f;
    return te >= -shift ? tx - te : tx + shift;
}

// Chord of a crystal whose semi-axes are scaled by s (about its centre) and the chord's
// derivative with respect to s.  The chord is crystal_hit's on the scaled record; the
// derivative is the closed form in the crystal frame: for the box, the exit and (when
// the ray starts outside) entry planes move by h_k / |d_k| per unit s; for the
// ellipsoid, with P = l / h and Q = d / h in the unit frame, the roots are
// (-P.Q -+ sqrt((P.Q)^2 - Q.Q (P.P - s^2))) / Q.Q, so each root the chord uses moves
// by s / sqrt(...).
FG_HD bool crystal_chord_scaled(const float c[3], const float q[4], const float h[3], uint32_t morph,
                                float s, float ox, float oy, float oz, float dx, float dy, float dz,
                                float &chord, float &dchord)
{
    const float hs[3] = {h[0] * s, h[1] * s, h[2] * s};
    float t0, t1;
    if (!crystal_hit(c, q, hs, morph, ox, oy, oz, dx, dy, dz, t0, t1)) return false;
    chord = fmaxf(t1 - t0, 0.0f);
    float lx, ly, lz, ex, ey, ez;
    qinv_rotate(q, ox - c[0], oy - c[1], oz - c[2], lx, ly, lz);
    qinv_rotate(q, dx, dy, dz, ex, ey, ez);
    if (morph == 0u) {
        const float l[3] = {lx, ly, lz}, e[3] = {ex, ey, ez};
        float tmin = -INFINITY, tmax = INFINITY, gmin = 0.0f, gmax = 0.0f;
        for (int k = 0; k < 3; k++) {
            const float inv = 1.0f / e[k];
            const float a1 = (-hs[k] - l[k]) * inv, a2 = (hs[k] - l[k]) * inv;
            const float lo = fminf(a1, a2), hi = fmaxf(a1, a2);
            const float g = h[k] * fabsf(inv);
            if (lo > tmin) { tmin = lo; gmin = g; }
            if (hi < tmax) { tmax = hi; gmax = g; }
        }
        dchord = gmax + (tmin > 0.0f ? gmin : 0.0f);
        if (!(dchord < INFINITY)) dchord = 0.0f;
        return true;
    }
    const float Px = lx / h[0], Py = ly / h[1], Pz = lz / h[2];
    const float Qx = ex / h[0], Qy = ey / h[1], Qz = ez / h[2];
    const float QQ = Qx * Qx + Qy * Qy + Qz * Qz, PQ = Px * Qx + Py * Qy + Pz * Qz;
    const float PP = Px * Px + Py * Py + Pz * Pz;
    const float r = sqrtf(fmaxf(PQ * PQ - QQ * (PP - s * s), 0.0f));
    const float enter = (-PQ - r) / QQ;
    dchord = r > 0.0f ? (enter > 0.0f ? 2.0f : 1.0f) * s / r : 0.0f;
    return true;
}

// AABB::intersect_ray (bvh.rs:96-124)
FG_HD bool box_hit(const float lo[3], const float hi[3], float ox, float oy, float oz,
                   float ix, float iy, float iz)
{
    const float t1x = (lo[0] - ox) * ix, t2x = (hi[0] - ox) * ix;
    const float tminx = fminf(t1x, t2x), tmaxx = fmaxf(t1x, t2x);
    const float t1y = (lo[1] - oy) * iy, t2y = (hi[1] - oy) * iy;
    const float tminy = fminf(t1y, t2y), tmaxy = fmaxf(t1y, t2y);
    const float t1z = (lo[2] - oz) * iz, t2z = (hi[2] - oz) * iz;
    const float tminz = fminf(t1z, t2z), tmaxz = fmaxf(t1z, t2z);
    const float te = fmaxf(fmaxf(tminx, tminy), tminz);
    const float tx = fminf(fminf(tmaxx, tmaxy), tmaxz);
    return tx >= te && tx > 0.0f;
}

}  // namespace fg
