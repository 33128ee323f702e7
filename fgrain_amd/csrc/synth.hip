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

// ------------------------------------------------- compaction and crystal attributes
constexpr double kTauD = 6.283185307179586, kPiD = 3.141592653589793;

// The per-point stream of device synthesis: ctr = (lo32(point), hi32(point), block,
// kTagAttr), words consumed in order (as the host form in host.cpp draws them).
struct PointStream {
    uint32_t k0, k1, p0, p1, block = 0, buf[4];
    int pos = 4;
    __device__ PointStream(uint64_t seed, uint64_t p)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), p0((uint32_t)p), p1((uint32_t)(p >> 32)) {}
    __device__ uint32_t u32()
    {
        if (pos == 4) {
            buf[0] = p0;
            buf[1] = p1;
            buf[2] = block++;
            buf[3] = kTagAttr;
            philox(buf, k0, k1);
            pos = 0;
        }
        return buf[pos++];
    }
    __device__ double uniform()
    {
        const uint64_t lo = u32(), hi = u32();
        return (double)((lo | (hi << 32)) >> 11) * 0x1p-53;
    }
    __device__ double normal(double mean, double sd)
    {
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return mean + sd * (sqrt(-2.0 * log(u1)) * cos(kTauD * u2));
    }
};

// emulsion.rs:160-207 for one point: radius, sensitivity, orientation, semi-axes
__device__ void point_attributes(const SynthAttrs &A, double rmax, uint64_t seed, uint64_t p, double *d,
                                 double *q, double &sens)
{
    PointStream rng(seed, p);
    double rad = exp(rng.normal(A.mu, A.size_sd));
    rad = fmin(fmax(rad, A.min_radius), rmax);
    sens = fmax(rng.normal(A.base_sens, A.sens_sd), 0.1);
    if (A.morph == 0u || A.morph == 2u) {
        const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        // UnitQuaternion::from_euler_angles(2 pi u1, pi u2, 2 pi u3)
        const double ro = u1 * kTauD * 0.5, pi_ = u2 * kPiD * 0.5, ya = u3 * kTauD * 0.5;
        const double sr = sin(ro), cr = cos(ro), sp = sin(pi_), cp = cos(pi_), sy = sin(ya), cy = cos(ya);
        q[0] = cr * cp * cy + sr * sp * sy;
        q[1] = sr * cp * cy - cr * sp * sy;
        q[2] = cr * sp * cy + sr * cp * sy;
        q[3] = cr * cp * sy - sr * sp * cy;
        d[0] = d[1] = d[2] = rad;
    } else {
        const double ht = fmax(rad / A.aspect, A.min_half_thickness);
        const double tilt = -0.17 + 0.34 * rng.uniform();
        double ax = -1.0 + 2.0 * rng.uniform(), ay = -1.0 + 2.0 * rng.uniform(), az = 0.0;
        const double nn = sqrt(ax * ax + ay * ay + az * az);
        ax /= nn;
        ay /= nn;
        az /= nn;
        const double sh = sin(tilt * 0.5), ch = cos(tilt * 0.5);
        q[0] = ch;
        q[1] = ax * sh;
        q[2] = ay * sh;
        q[3] = az * sh;
        d[0] = rad;
        d[1] = rad;
        d[2] = ht;
    }
}

__global__ void __launch_bounds__(256) synth_count_kernel(const float4 *cells, const SynthTile *tiles, uint32_t *counts)
{
    __shared__ uint32_t wc[4];
    const SynthTile T = tiles[blockIdx.x];
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < T.n_cells; i += 256) c += cells[T.cell_off + i].w != 0.0f ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o, 64);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ void __launch_bounds__(256) synth_fill_kernel(const float4 *cells, const SynthTile *tiles, SynthAttrs A,
                                                         double *pos, double *dims, double *quat, double *sens)
{
    __shared__ uint32_t wc[4];
    const SynthTile T = tiles[blockIdx.x];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t running = 0;
    for (uint32_t c0 = 0; c0 < T.n_cells; c0 += 256) {
        const uint32_t i = c0 + threadIdx.x;
        const float4 cl = i < T.n_cells ? cells[T.cell_off + i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const bool has = cl.w != 0.0f;
        const unsigned long long m = __ballot(has);
        if (lane == 0) wc[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, tot = 0;
        for (uint32_t w = 0; w < 4; w++) {
            if (w < wave) before += wc[w];
            tot += wc[w];
        }
        if (has) {
            const uint32_t k = running + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            const size_t sl = (size_t)T.crystal_base + k;
            pos[3 * sl] = (double)cl.x;
            pos[3 * sl + 1] = (double)cl.y;
            pos[3 * sl + 2] = (double)cl.z;
            double sv;
            point_attributes(A, T.rmax, T.seed, k, &dims[3 * sl], &quat[4 * sl], sv);
            sens[sl] = sv;
        }
        running += tot;
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_synth_count(const float4 *cells, const SynthTile *tiles, uint32_t n_tiles, uint32_t *counts,
                              hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(synth_count_kernel, dim3(n_tiles), dim3(256), 0, s, cells, tiles, counts);
    return hipGetLastError();
}

hipError_t launch_synth_fill(const float4 *cells, const SynthTile *tiles, uint32_t n_tiles, const SynthAttrs &A,
                             double *pos, double *dims, double *quat, double *sens, hipStream_t s)
{
    (void)hipGetLastError();  // report only this launcher's launches
    if (n_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(synth_fill_kernel, dim3(n_tiles), dim3(256), 0, s, cells, tiles, A, pos, dims, quat, sens);
    return hipGetLastError();
}

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
