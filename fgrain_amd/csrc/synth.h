// Device Poisson-disk dart throwing (synth.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

struct DartArgs {
    float4 *cells;            // gx * gy * gz, x fastest; w = 1 when the cell holds a point
    uint32_t gx, gy, gz;
    float cell, r2, W, H, D;  // cell side r / sqrt(3), squared minimum distance, extents
    uint32_t key0, key1;
};

// `rounds` rounds over the 27 phases for n_tiles grids at once (device array of DartArgs,
// each with its own zeroed cells); max_phase_cells bounds the cells of one phase in any
// grid (ceil(g / 3) per axis).
hipError_t launch_darts_tiles(const DartArgs *tiles, uint32_t n_tiles, uint32_t max_phase_cells,
                              uint32_t rounds, hipStream_t s);

// One tile of a device-synthesised frame: its dart grid and where its crystals go.
struct SynthTile {
    uint64_t cell_off;        // first cell of the tile's grid in the cells buffer
    uint32_t n_cells;
    uint32_t crystal_base;    // first crystal slot of the tile (set after counting)
    uint64_t seed;            // the tile's synthesis seed (per-point Philox streams)
    double rmax;              // largest radius, r_min * 0.48 (emulsion.rs:166)
};

// Crystal attribute parameters of emulsion.rs:160-207 shared by the tiles.
struct SynthAttrs {
    double mu;                // ln(mean crystal radius)
    double size_sd, min_radius, base_sens, sens_sd, aspect, min_half_thickness;
    uint32_t morph;
    uint32_t pad;
};

// counts[t] = points in tile t's grid (one block per tile).
hipError_t launch_synth_count(const float4 *cells, const SynthTile *tiles, uint32_t n_tiles, uint32_t *counts,
                              hipStream_t s);
// Tile t's points in cell order into slots crystal_base.. (positions as f64), each with
// the attributes of emulsion.rs:160-207 drawn from its own Philox stream (seed, index in
// the tile): dims 3, quat 4 (w, i, j, k), sens 1 per slot.
hipError_t launch_synth_fill(const float4 *cells, const SynthTile *tiles, uint32_t n_tiles, const SynthAttrs &A,
                             double *pos, double *dims, double *quat, double *sens, hipStream_t s);

}  // namespace fg
