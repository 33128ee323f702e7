// Batched frame pipeline: every tile of a frame in one pass of device kernels
// (transport -> development -> developed-crystal BVH -> scan), replayed as a hipGraph.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "kernels.h"

namespace fg {

struct DevelopParams {
    uint32_t threshold;
    uint32_t pad;
    double fe, pf;
};

struct ScanCommon {
    uint32_t g, ss, mode, rays_per_lane;
    float z0, sigma;
};

struct BatchedFrame;

struct BatchedInput {
    const Node *nodes;              // transport nodes of all tiles
    const Crystal *recs;            // transport records of all tiles
    uint32_t *counts;               // latent counts per slot
    const float *emap;              // frame exposure map
    float *dneg;                    // frame output
    unsigned long long *stats;      // transport counters (5)
    uint32_t frame_w, frame_h;
    int64_t n_crystals;
    std::vector<TileDev> tiles;
    // host arrays in slot order (per tile: slot k holds crystal bvh.idx[k])
    std::vector<double> pos64, dims64;
    std::vector<uint32_t> orig;
    // or, for crystals already on the device (crystal order, tile t's from its
    // crystal_base), their attributes and each slot's crystal index within its tile;
    // then the tiles' maxb are taken as given (an upper bound of their largest semi-axis)
    const double *dev_pos = nullptr, *dev_dims = nullptr;
    const uint32_t *dev_leaf = nullptr;
    TransportArgs transport;        // optics and shared fields; per-tile fields overwritten
    uint32_t transport_waves = 6;   // occupancy of the transport kernel: 6 or 7 waves per SIMD
    DevelopParams develop;
    ScanCommon scan;
};

hipError_t batched_create(const BatchedInput &in, hipStream_t s, BatchedFrame **out);
// Render the frame; use_graph replays the captured hipGraphs of its three stages
// (transport; development + developed-crystal BVH; scan), captured on first use.
hipError_t batched_render(BatchedFrame *b, hipStream_t s, bool use_graph);
// Development and the developed-crystal BVH of the latent counts already in the input's
// counts (direct launches; batched_kernel_ms then reports this stage alone).
hipError_t batched_develop(BatchedFrame *b, hipStream_t s);
// The graph call that failed when batched_render with use_graph returned an error.
const char *batched_graph_failure(const BatchedFrame *b);
hipError_t batched_read_tiles(BatchedFrame *b, hipStream_t s, std::vector<TileDev> &tiles,
                              uint32_t *err_flags);
hipError_t batched_read_develop(BatchedFrame *b, hipStream_t s, std::vector<uint8_t> &dev,
                                std::vector<double> &fe, std::vector<int32_t> &clump);
// Kernel times of the last render (HIP events around transport, development + BVH,
// and scan, recorded on the render stream between the stages).
hipError_t batched_kernel_ms(BatchedFrame *b, float *transport, float *develop, float *scan);
void batched_destroy(BatchedFrame *b);

// Transport BVHs (BVH::build, bvh.rs:157-295) of many tiles' crystals on the device, with
// the batched frame's level kernels: tile t's crystals are slots base[t] .. base[t] +
// count[t] in crystal order (pos, dims 3 per slot; quat 4; sens 1; one morphology).
// Writes tile t's threaded nodes from nodes + 2 base[t] (node_count[t] of them), its
// records in leaf order from recs + base[t], and leaf_slot[base[t] + k] = the index
// within the tile of the crystal at leaf position k.  err: [0] a node deeper than the reference's stack, [1] a leaf of more
// than 15 crystals, [2] an unfinished build.  Synchronises s.
struct TransportBuildIn {
    const double *pos, *dims, *quat, *sens;
    uint32_t morph;
    int64_t n;
    std::vector<uint32_t> base, count;
};
hipError_t build_transport_bvhs(const TransportBuildIn &in, Node *nodes, Crystal *recs, uint32_t *leaf_slot,
                                std::vector<uint32_t> &node_count, uint32_t err[4], hipStream_t s);

}  // namespace fg
