// gfx950 kernels of the photon-transport + scan path.
//
// Traversal everywhere is the stackless threaded walk of fg_layout.h: one thread
// follows i -> i + 1 on an interior hit and i -> skip otherwise, which visits nodes in
// the reference's depth-first order (bvh.rs:300-347, right child first) and therefore
// reports hits in the reference's callback order, without a per-lane stack in LDS or
// scratch.  Nodes are read as two 16-byte loads, crystal records as three.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace fg {

namespace {

using namespace dev;

// ------------------------------------------------------------------- transport ---
// One thread per launched packet (plain per-tile launch).
__global__ void __launch_bounds__(256) transport_kernel(TransportArgs A)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    PhotonCounters cnt;
    if (p < A.n_photons) run_photon(A, p, cnt);
    if (A.stats) flush_counters(cnt, A.stats);
}

// Persistent transport of one tile (kPre: deposits pre-reduced across the wave).
template <bool kPre>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 6))) transport_persistent_kernel(TransportArgs A, unsigned long long *work)
{
    LaneCounters cnt;
    transport_persistent<false, kCounts, kPre>(A, nullptr, 0u, nullptr, nullptr, nullptr, nullptr, work, A.n_photons, cnt);
    if (A.stats) flush_lane_counters(cnt, A.stats);
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
    const uint32_t L = scan_lanes(A.g * A.g * A.ss * A.ss, A.rays_per_lane);
    const uint32_t pix = blockIdx.x * (256u / L) + threadIdx.x / L;
    const bool valid = pix < A.w_px * A.h_px;
    const uint32_t px = valid ? pix % A.w_px : 0u, py = valid ? pix / A.w_px : 0u;
    const float T = scan_pixel(valid, pix, px, py, threadIdx.x % L, L, A.g, A.ss, A.pw, A.ph, A.z0,
                               A.sigma, A.key0, A.key1, A.nodes, A.qnodes, A.n_nodes, A.recs, A.consts);
    if (valid && threadIdx.x % L == 0u) A.out[(size_t)py * A.ostride + px] = scan_value(T, A.mode);
}

__global__ void __launch_bounds__(256) down_constants_kernel(float4 *recs, const float4 *lo, uint32_t n,
                                                             float4 *consts)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Rec r = load_rec(recs, k);
    float c[4];
    down_constants(r.q, r.h, r.morph, c);
    consts[2 * k] = make_float4(c[0], c[1], c[2], c[3]);
    consts[2 * k + 1] = lo[k];
    recs[3 * k] = make_float4(r.c[0], r.c[1], r.c[2], c[3]);   // radius in the sensitivity word
}

__global__ void __launch_bounds__(256) scan_qnodes_kernel(const Node *nodes, uint32_t n, ScanQNode *qnodes)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float bx, by, sc;
    scan_qframe(nodes[0], bx, by, sc);
    qnodes[i] = scan_qnode(nodes[i], bx, by, sc);
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

// ------------------------------------------------- VALU counter calibration -----
// A known instruction count for the profiler's SQ_INSTS_VALU: per wave, 64 v_fmac_f32
// per iteration plus 4 outside the loop (v_cvt, two v_mov, v_cmp; the store is never
// taken), no other VALU instruction (checked in the gfx950 assembly of this kernel).
#define FG_FMAC8(x)                                                                              \
    asm volatile("v_fmac_f32 %0, %1, %2\n v_fmac_f32 %0, %1, %2\n v_fmac_f32 %0, %1, %2\n"        \
                 " v_fmac_f32 %0, %1, %2\n v_fmac_f32 %0, %1, %2\n v_fmac_f32 %0, %1, %2\n"       \
                 " v_fmac_f32 %0, %1, %2\n v_fmac_f32 %0, %1, %2"                                  \
                 : "+v"(x)                                                                       \
                 : "v"(b), "v"(c));
__global__ void __launch_bounds__(256) valu_probe_kernel(float *out, uint32_t iters, float b, float c)
{
    float a = (float)threadIdx.x;
    for (uint32_t i = 0; i < iters; i++) {
        FG_FMAC8(a) FG_FMAC8(a) FG_FMAC8(a) FG_FMAC8(a) FG_FMAC8(a) FG_FMAC8(a) FG_FMAC8(a) FG_FMAC8(a)
    }
    if (a == 1.2345f) out[blockIdx.x] = a;
}
#undef FG_FMAC8

}  // namespace

hipError_t launch_valu_probe(float *out, uint32_t blocks, uint32_t iters, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (blocks == 0 || iters == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, s, out, iters, 1.0f, 0.0f);
    return hipGetLastError();
}

hipError_t launch_scan_qnodes(const Node *nodes, uint32_t n_nodes, ScanQNode *qnodes, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n_nodes == 0) return hipSuccess;
    hipLaunchKernelGGL(scan_qnodes_kernel, dim3((n_nodes + 255) / 256), dim3(256), 0, s, nodes, n_nodes, qnodes);
    return hipGetLastError();
}

hipError_t launch_transport(const TransportArgs &a, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (a.n_photons == 0) return hipSuccess;
    const uint64_t blocks = (a.n_photons + 255) / 256;
    hipLaunchKernelGGL(transport_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_down_constants(float4 *recs, const float4 *lo, uint32_t n, float4 *consts, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(down_constants_kernel, dim3((n + 255) / 256), dim3(256), 0, s, recs, lo, n, consts);
    return hipGetLastError();
}

hipError_t launch_transport_persistent(const TransportArgs &a, unsigned long long *work, int grid,
                                       hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (a.n_photons == 0) return hipSuccess;
    if (a.prereduce)
        hipLaunchKernelGGL(transport_persistent_kernel<true>, dim3(grid), dim3(256), 0, s, a, work);
    else
        hipLaunchKernelGGL(transport_persistent_kernel<false>, dim3(grid), dim3(256), 0, s, a, work);
    return hipGetLastError();
}

int transport_persistent_blocks_per_cu()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(transport_persistent_kernel<false>),
                                                     256, 0) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

hipError_t launch_scan(const ScanArgs &a, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (a.w_px == 0 || a.h_px == 0) return hipSuccess;
    const uint32_t ppb = 256u / scan_lanes(a.g * a.g * a.ss * a.ss, a.rays_per_lane);
    const uint32_t blocks = (uint32_t)(((uint64_t)a.w_px * a.h_px + ppb - 1) / ppb);
    hipLaunchKernelGGL(scan_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_all_hits(const Node *nodes, uint32_t n_nodes, const float4 *recs,
                           const float *o, const float *d, int64_t n_rays, int32_t max_hits,
                           int32_t *n_hits, int32_t *slot, float *t0, float *t1, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n_rays == 0) return hipSuccess;
    hipLaunchKernelGGL(all_hits_kernel, dim3((uint32_t)((n_rays + 255) / 256)), dim3(256), 0, s,
                       nodes, n_nodes, recs, o, d, n_rays, max_hits, n_hits, slot, t0, t1);
    return hipGetLastError();
}

hipError_t launch_first_hit(const Node *nodes, uint32_t n_nodes, const float4 *recs,
                            const float *o, const float *d, int64_t n_rays, int32_t *slot,
                            float *t0, float *t1, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n_rays == 0) return hipSuccess;
    hipLaunchKernelGGL(first_hit_kernel, dim3((uint32_t)((n_rays + 255) / 256)), dim3(256), 0, s,
                       nodes, n_nodes, recs, o, d, n_rays, slot, t0, t1);
    return hipGetLastError();
}

// ------------------------------------------------------- scan high-pass ---------
// delta_d = D - G(D) of the scan-only compatibility entry (gaussian.wgsl:21-59): the
// shader's 17-tap blur with its literal weights (labelled sigma 2.5; they sum to 1.146,
// so a flat field maps to -0.313 D) and clamp-to-edge taps, horizontal pass then
// vertical pass, each summed in tap order as one f32 product and one f32 addition per
// tap.  One 64 x 16 output tile per block: the clamped source tile
// with an 8-pixel halo (32 x 80) and the horizontal pass of its 32 rows stay in LDS, so
// D is read once and delta_d written once.
namespace {
constexpr int kHpR = 8, kHpX = 64, kHpY = 16;
__constant__ float kHpW[2 * kHpR + 1] = {0.002621f, 0.007802f, 0.019445f, 0.040523f, 0.070624f, 0.103073f,
                                         0.125925f, 0.134267f, 0.137440f, 0.134267f, 0.125925f, 0.103073f,
                                         0.070624f, 0.040523f, 0.019445f, 0.007802f, 0.002621f};
}  // namespace

__global__ void __launch_bounds__(256) highpass17_kernel(const float *in, float *out, uint32_t w, uint32_t h)
{
    __shared__ float src[kHpY + 2 * kHpR][kHpX + 2 * kHpR + 1];
    __shared__ float hor[kHpY + 2 * kHpR][kHpX + 1];
    const int x0 = (int)blockIdx.x * kHpX, y0 = (int)blockIdx.y * kHpY;
    // LDS cell (r, c) holds D at the clamped pixel (y0 - 8 + r, x0 - 8 + c), so the tap
    // k of output column i is cell i + k + 8 in both directions
    for (int k = threadIdx.x; k < (kHpY + 2 * kHpR) * (kHpX + 2 * kHpR); k += 256) {
        const int r = k / (kHpX + 2 * kHpR), c = k % (kHpX + 2 * kHpR);
        const int gx = min(max(x0 - kHpR + c, 0), (int)w - 1);
        const int gy = min(max(y0 - kHpR + r, 0), (int)h - 1);
        src[r][c] = in[(size_t)gy * w + gx];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < (kHpY + 2 * kHpR) * kHpX; k += 256) {
        const int r = k / kHpX, c = k % kHpX;
        float sum = 0.0f;
        for (int j = 0; j <= 2 * kHpR; j++) sum += src[r][c + j] * kHpW[j];
        hor[r][c] = sum;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kHpY * kHpX; k += 256) {
        const int r = k / kHpX, c = k % kHpX;
        const int gx = x0 + c, gy = y0 + r;
        if (gx >= (int)w || gy >= (int)h) continue;
        float sum = 0.0f;
        for (int j = 0; j <= 2 * kHpR; j++) sum += hor[r + j][c] * kHpW[j];
        out[(size_t)gy * w + gx] = src[r + kHpR][c + kHpR] - sum;
    }
}

hipError_t launch_highpass17(const float *in, float *out, uint32_t w, uint32_t h, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (w == 0 || h == 0) return hipSuccess;
    hipLaunchKernelGGL(highpass17_kernel, dim3((w + kHpX - 1) / kHpX, (h + kHpY - 1) / kHpY), dim3(256), 0, s,
                       in, out, w, h);
    return hipGetLastError();
}

hipError_t launch_otf3(const float *in, float *out, uint32_t w, uint32_t h, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (w == 0 || h == 0) return hipSuccess;
    hipLaunchKernelGGL(otf3_kernel, dim3((w + 15) / 16, (h + 15) / 16), dim3(256), 0, s, in, out, w, h);
    return hipGetLastError();
}

}  // namespace fg
