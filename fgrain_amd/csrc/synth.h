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

}  // namespace fg
