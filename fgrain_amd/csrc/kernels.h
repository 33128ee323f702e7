// Launch interfaces of the gfx950 kernels (kernels.hip).  Host code passes device
// pointers and plain parameter blocks; every launcher returns the hipError_t of its
// launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fg_layout.h"

namespace fg {

struct TransportArgs {
    const Node *nodes;
    const float4 *recs;      // 3 float4 per crystal record, leaf order
    const PreRec *pre;       // pre-test record per crystal (the persistent replay kernels)
    uint32_t n_nodes;
    uint32_t chunk;          // packets a wave takes from the work counter at a time (>= 64)
    const float *emap;       // exposure map (may be null: weight 1)
    uint32_t tw, th, estride;
    uint32_t ppp;            // > 0: stratified, photon p in pixel p / ppp
    uint32_t *counts;        // latent counts, leaf order
    unsigned long long *stats;  // bounces, nodes, prims, abs_events, abs_total
    float W, H, D, pw, ph;
    float mu_t, albedo, r0, hal_eff, mu_abs, theta_max, g, step_cap;
    uint32_t max_bounces;
    uint32_t key0, key1;
    uint64_t n_photons;
    // Exact pruning: a crystal point on the step segment lies within `reach` of the
    // crystal's AABB (corners of a rotated box stick out by at most (sqrt(3) - 1) times
    // its largest semi-axis; ellipsoids stay inside).  Nodes whose box, grown by reach,
    // misses the segment's bounding box hold no crystal that can absorb in this step.
    // reach < 0 disables the test (full infinite-ray traversal as the reference).
    float reach;
    // persistent transport: node steps continue until this percentage of the lanes
    // that are walking the BVH hold a pending leaf (100: all of them)
    uint32_t wait_pct;
    // film base below the emulsion (0 = the reference's in-place reflection): a
    // reflected packet re-enters displaced by base * (tan of the way down + tan of the
    // way up), and is lost when that leaves the patch
    float base;
    // 1: deposits are pre-reduced across the wave before the global atomics
    // (wave_aggregated_add); 0: one atomic per absorbing hit.  Same counts either way.
    uint32_t prereduce;
};

// Path replay of the expected model (adjoint.hip): the transport's packets without the
// rounding and roulette draws, depositing expected dose or accumulating gradients.
struct ReplayArgs {
    TransportArgs t;       // geometry, optics, exposure map, keys (t.counts unused)
    double *dose;          // dose pass: expected dose accumulators, transport slot order
    const float *gdose;    // gradient pass: dL/d(dose) per transport slot
    double *gmap;          // gradient pass: dL/dE accumulators (tw * th)
    double *partials;      // gradient pass: per block (mu_abs, g, R) sums, 4 doubles each
    unsigned long long *stats;  // packets, bounces, events
    float hal_unit;        // 10^(-2 OD): t.hal_eff = R * hal_unit
    float refl;            // antihalation reflectance R
    float cut;             // a packet ends once its relative weight and R tangent are below
    float pad;
};

// Per-tile descriptor shared by host and device.
struct TileDev {
    uint32_t x0, y0, tw, th;        // frame placement (pixels)
    float W, H, pw, ph;             // extents and pixel size (length units)
    uint32_t key0, key1;            // transport seed (tile seed + 1)
    uint32_t skey0, skey1;          // scan seed (tile seed + 2)
    uint32_t node_base, n_nodes;    // transport BVH nodes
    uint32_t crystal_base, n;       // crystal slots (leaf order)
    uint64_t item_base;             // first 64-packet work item
    uint64_t n_photons;
    uint64_t photon_base;           // first packet id of the tile in the frame
    uint32_t block_base, blocks_x;  // scan blocks of 16 x 16 pixels
    uint32_t blocks, out_stride;    // scan output row stride
    uint64_t out_base;              // first output value of the tile
    double maxb;                    // largest semi-axis among the tile's crystals
    // results of the last render
    uint32_t n_dev, clumps, max_clump, dev_nodes;
};

struct ScanArgs {
    const Node *nodes;
    const ScanQNode *qnodes;   // the scan's grid nodes of nodes (launch_scan_qnodes)
    const float4 *recs;
    const float4 *consts;    // 2 float4 per record (launch_down_constants)
    uint32_t n_nodes;
    uint32_t w_px, h_px;     // output grid
    uint32_t g, ss;          // strata per side, apertures per side
    uint32_t mode;
    float pw, ph, z0;        // pixel size, ray origin height (D + 0.1)
    float sigma;
    uint32_t key0, key1;
    float *out;              // w_px * h_px
    uint32_t ostride;        // row stride of out
    uint32_t rays_per_lane;  // scan shape (scan_lanes); 0 or 1: one lane per ray
};

hipError_t launch_transport(const TransportArgs &a, hipStream_t s);
// consts[2 k] = down_constants of record k (3 float4 per record), consts[2 k + 1] = lo[k],
// the residual of its centre's rounding to f32 (x, y, z, 0); k < n
hipError_t launch_down_constants(float4 *recs, const float4 *lo, uint32_t n, float4 *consts, hipStream_t s);
// qnodes[i] = scan_qnode(nodes[i]) on the grid of nodes[0] (scan_qframe), i < n_nodes
hipError_t launch_scan_qnodes(const Node *nodes, uint32_t n_nodes, ScanQNode *qnodes, hipStream_t s);
// Persistent transport of one tile (packets 0 .. a.n_photons - 1); work must point to a
// zeroed 64-bit counter; grid = resident blocks of the device.
hipError_t launch_transport_persistent(const TransportArgs &a, unsigned long long *work, int grid,
                                       hipStream_t s);
int transport_persistent_blocks_per_cu();
hipError_t launch_scan(const ScanArgs &a, hipStream_t s);
hipError_t launch_all_hits(const Node *nodes, uint32_t n_nodes, const float4 *recs,
                           const float *o, const float *d, int64_t n_rays, int32_t max_hits,
                           int32_t *n_hits, int32_t *slot, float *t0, float *t1,
                           hipStream_t s);
hipError_t launch_first_hit(const Node *nodes, uint32_t n_nodes, const float4 *recs,
                            const float *o, const float *d, int64_t n_rays, int32_t *slot,
                            float *t0, float *t1, hipStream_t s);
// blocks x 256 threads of the VALU calibration kernel (64 * iters + 4 VALU per wave);
// out: blocks floats (never written in practice)
hipError_t launch_valu_probe(float *out, uint32_t blocks, uint32_t iters, hipStream_t s);
hipError_t launch_otf3(const float *in, float *out, uint32_t w, uint32_t h, hipStream_t s);
// delta_d = D - the 17-tap blur of D, clamp-to-edge (gaussian.wgsl:14-59); in != out
hipError_t launch_highpass17(const float *in, float *out, uint32_t w, uint32_t h, hipStream_t s);

}  // namespace fg
