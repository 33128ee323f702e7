// Device Poisson-disk sampling for emulsion synthesis: parallel dart throwing on a
// grid of cell r / sqrt(3) (one point per cell at most, as the reference's Bridson grid,
// poisson.rs:14-132).  Cells are split into 27 phases by (i mod 3, j mod 3, k mod 3);
// two cells of one phase differ by at least 3 in some axis, so their points are at least
// 2 r / sqrt(3) > r apart and every cell of a phase can accept a candidate at once, its
// neighbour test (+-2 cells) reading only cells of other phases.  Each round visits the
// 27 phases; an empty cell draws one candidate per round from its own Philox stream.
#include <hip/hip_runtime.h>

#include "fg_layout.h"
#include "synth.h"

namespace fg {

namespace {

__device__ __forceinline__ void dart(const DartArgs &A, uint32_t i0, uint32_t j0, uint32_t k0, uint32_t round,
                                     uint32_t t)
{
    if (i0 >= A.gx || j0 >= A.gy || k0 >= A.gz) return;
    const uint32_t nx = (A.gx - i0 + 2u) / 3u, ny = (A.gy - j0 + 2u) / 3u, nz = (A.gz - k0 + 2u) / 3u;
    if (t >= nx * ny * nz) return;
    const uint32_t i = i0 + 3u * (t % nx), j = j0 + 3u * ((t / nx) % ny), k = k0 + 3u * (t / (nx * ny));
    const uint32_t cell = i + A.gx * (j + A.gy * k);
    if (A.cells[cell].w != 0.0f) return;
    uint32_t c[4] = {cell, round, kTagDart, 0u};
    philox(c, A.key0, A.key1);
    const float x = ((float)i + u24(c[0])) * A.cell;
    const float y = ((float)j + u24(c[1])) * A.cell;
    const float z = ((float)k + u24(c[2])) * A.cell;
    if (!(x < A.W && y < A.H && z < A.D)) return;
    const int zi = (int)k, yi = (int)j, xi = (int)i;
    for (int dz = -2; dz <= 2; dz++) {
        const int kk = zi + dz;
        if (kk < 0 || kk >= (int)A.gz) continue;
        for (int dy = -2; dy <= 2; dy++) {
            const int jj = yi + dy;
            if (jj < 0 || jj >= (int)A.gy) continue;
            for (int dx = -2; dx <= 2; dx++) {
                const int ii = xi + dx;
                if (ii < 0 || ii >= (int)A.gx) continue;
                const float4 q = A.cells[(uint32_t)ii + A.gx * ((uint32_t)jj + A.gy * (uint32_t)kk)];
                if (q.w == 0.0f) continue;
                const float ex = q.x - x, ey = q.y - y, ez = q.z - z;
                if ((ex * ex + ey * ey) + ez * ez < A.r2) return;
            }
        }
    }
    A.cells[cell] = make_float4(x, y, z, 1.0f);
}

// every tile of a frame at once: blockIdx.y selects the tile
__global__ void __launch_bounds__(256) darts_tiles_kernel(const DartArgs *tiles, uint32_t i0, uint32_t j0,
                                                          uint32_t k0, uint32_t round)
{
    const DartArgs A = tiles[blockIdx.y];
    dart(A, i0, j0, k0, round, blockIdx.x * blockDim.x + threadIdx.x);
}

}  // namespace

hipError_t launch_darts_tiles(const DartArgs *tiles, uint32_t n_tiles, uint32_t max_phase_cells,
                              uint32_t rounds, hipStream_t s)
{
    if (n_tiles == 0 || max_phase_cells == 0) return hipSuccess;
    const dim3 grid((max_phase_cells + 255u) / 256u, n_tiles);
    for (uint32_t r = 0; r < rounds; r++)
        for (uint32_t k0 = 0; k0 < 3u; k0++)
            for (uint32_t j0 = 0; j0 < 3u; j0++)
                for (uint32_t i0 = 0; i0 < 3u; i0++) {
                    darts_tiles_kernel<<<grid, 256, 0, s>>>(tiles, i0, j0, k0, r);
                    hipError_t e = hipGetLastError();
                    if (e != hipSuccess) return e;
                }
    return hipSuccess;
}

}  // namespace fg
