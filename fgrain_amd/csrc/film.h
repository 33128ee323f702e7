// Full-frame black-and-white compositing on the GPU (image_pipeline.rs:133-891): the
// layout of overlapping tiles, the optics and compositing kernels' launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fg {

// Overlapping tile layout of the reference's tiled renderer (image_pipeline.rs:380-421)
// and the Hann ramps of its blend (:614-636).  Tile (tx, ty) covers columns
// [xs[tx], xs[tx] + tw[tx]) and rows [ys[ty], ys[ty] + th[ty]); its D_neg values start at
// yoff[ty] * wext + th[ty] * xoff[tx] in the concatenated tile buffer (tiles in frame
// order, each row-major).  ramp[k] (k < 2 pad) is the rising weight 0.5 - 0.5 cos(pi k /
// 2 pad) and ramp[2 pad + k] the falling one 0.5 + 0.5 cos(pi k / 2 pad), evaluated on
// the host with the reference's f32 operations.
struct FilmLayout {
    uint32_t W, H, step, pad, ntx, nty;
    uint64_t wext;                  // sum of the tile widths of one tile row
    const uint32_t *xs, *tw, *xoff; // per tile column
    const uint32_t *ys, *th, *yoff; // per tile row
    const float *ramp;              // 4 pad entries
};

// Compositing parameters (image_pipeline.rs:651-731).
struct FilmParams {
    float strength;       // effective grain strength
    float base_amp;       // base_nominal * (1 + 0.15 push)
    float push;           // effective push (tone curve when > 0)
    uint32_t otf;         // scanner OTF on the grain field
    uint32_t rad;         // Gaussian radius
    uint32_t pad_;
    const float *kern;    // 2 rad + 1 normalised Gaussian weights
};

// Both structs are kernel arguments built by host code and read by device code, which
// are compiled apart: their layout is pinned here for both compiles.
static_assert(sizeof(FilmLayout) == 88 && offsetof(FilmLayout, wext) == 24 && offsetof(FilmLayout, xs) == 32 &&
                  offsetof(FilmLayout, ramp) == 80,
              "FilmLayout layout");
static_assert(sizeof(FilmParams) == 32 && offsetof(FilmParams, rad) == 16 && offsetof(FilmParams, kern) == 24,
              "FilmParams layout");

// The Hann blend kernel belongs to the compositing kernels that are switched off (a
// device memory fault on their first launch, cause not yet found); FG_FILM_COMPOSITE=1
// at compile time switches it on.
#ifndef FG_FILM_COMPOSITE
#define FG_FILM_COMPOSITE 0
#endif
constexpr bool kFilmBlendEnabled = FG_FILM_COMPOSITE != 0;

// span of tile t along one axis (image_pipeline.rs:400-410)
inline void film_span(uint32_t t, uint32_t step, uint32_t pad, uint32_t size, uint32_t &start, uint32_t &len)
{
    const uint64_t x0 = (uint64_t)t * step;
    const uint64_t s = x0 > pad ? x0 - pad : 0;
    const uint64_t e = x0 + step + pad < size ? x0 + step + pad : size;
    start = (uint32_t)s;
    len = (uint32_t)(e - s);
}

// bounding box of the nonzero pixels (first row, last row, first column, last column)
// into bbox[0..3] and the optics into h_eff (image_pipeline.rs:137-249); bbox holds
// film_bbox_size(W, H) ints (the per-block boxes follow the first four); W * H <= 2^31
size_t film_bbox_size(uint32_t W, uint32_t H);
hipError_t launch_film_optics(const uint8_t *gray, uint32_t W, uint32_t H, int32_t *bbox, float *h_eff,
                              hipStream_t s);
// convert_to_panchromatic_luma of interleaved RGB (image_pipeline.rs:284-293)
hipError_t launch_film_luma(const uint8_t *rgb, uint64_t n, float rw, float gw, float bw, uint8_t *gray,
                            hipStream_t s);
// d_norm = Hann blend of the tiles (image_pipeline.rs:593-654)
hipError_t launch_film_blend(const FilmLayout &L, const float *tiles, float *d_norm, hipStream_t s);
// blend, low-frequency separation, statistics, grain, OTF and the 8-bit output.  tmp and
// d_high: W x H f32 work buffers; partials: film_partials_size(W, H) doubles; fstats: 4
// floats (high mean, high std, norm factor); grain_out optional.
size_t film_partials_size(uint32_t W, uint32_t H);
hipError_t launch_film_composite(const FilmLayout &L, const FilmParams &P, const uint8_t *gray,
                                 const float *h_eff, const float *tiles, float *d_norm, float *tmp,
                                 float *d_high, double *partials, float *fstats, float *grain_out,
                                 uint8_t *out, hipStream_t s);

}  // namespace fg
