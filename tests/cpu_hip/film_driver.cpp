// Drives the host build of the film kernels: reads gray, h_effective, tiles and the
// plan from a file, runs the optics and the compositing launch sequence of
// launch_film_optics / launch_film_composite, writes h_effective, d_norm, d_high, grain,
// the 8-bit output and the statistics.
#include <cmath>
#include <vector>

#include "film.h"

using namespace fg;

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t hdr[6];   // W, H, step, pad, rad, otf
    float par[4];      // strength, base amplitude, push, Gaussian sigma
    if (fread(hdr, 4, 6, f) != 6 || fread(par, 4, 4, f) != 4) return 2;
    const uint32_t W = hdr[0], H = hdr[1], step = hdr[2], pad = hdr[3], rad = hdr[4], otf = hdr[5];
    const size_t n = (size_t)W * H;
    // the tables fg_film_composite builds on the host, with the same f32 operations
    std::vector<float> kern;
    float ksum = 0.0f;
    for (int i = -(int)rad; i <= (int)rad; i++) {
        const float q = (float)i / par[3];
        const float w = expf(-0.5f * (q * q));
        kern.push_back(w);
        ksum += w;
    }
    for (auto &w : kern) w /= fmaxf(ksum, 1e-6f);
    std::vector<uint8_t> gray(n);
    if (fread(gray.data(), 1, n, f) != n) return 2;
    const uint32_t ntx = (W + step - 1) / step, nty = (H + step - 1) / step;
    std::vector<uint32_t> xs(ntx), tw(ntx), xoff(ntx), ys(nty), th(nty), yoff(nty);
    uint64_t wext = 0, hext = 0;
    for (uint32_t t = 0; t < ntx; t++) { film_span(t, step, pad, W, xs[t], tw[t]); xoff[t] = (uint32_t)wext; wext += tw[t]; }
    for (uint32_t t = 0; t < nty; t++) { film_span(t, step, pad, H, ys[t], th[t]); yoff[t] = (uint32_t)hext; hext += th[t]; }
    std::vector<float> ramp(4 * pad + 1);
    for (uint32_t k = 0; k < 2 * pad; k++) {
        const float c = cosf(3.14159265358979323846f * ((float)k / (float)(2 * pad)));
        ramp[k] = 0.5f - 0.5f * c;
        ramp[2 * pad + k] = 0.5f + 0.5f * c;
    }
    std::vector<float> tiles(wext * hext);
    if (fread(tiles.data(), 4, tiles.size(), f) != tiles.size()) return 2;
    fclose(f);
    FilmLayout L{W, H, step, pad, ntx, nty, wext, xs.data(), tw.data(), xoff.data(), ys.data(), th.data(), yoff.data(),
                 ramp.data()};
    FilmParams P{par[0], par[1], par[2], otf, rad, 0u, kern.data()};
    std::vector<float> heff(n), dn(n), tmp(n), dh(n), grain(n), fs(4);
    std::vector<double> part(film_partials_size(W, H));
    std::vector<uint8_t> out(n);
    std::vector<int32_t> bbox(film_bbox_size(W, H));
    if (launch_film_optics(gray.data(), W, H, bbox.data(), heff.data(), nullptr)) return 3;
    if (launch_film_composite(L, P, gray.data(), heff.data(), tiles.data(), dn.data(), tmp.data(), dh.data(),
                              part.data(), fs.data(), grain.data(), out.data(), nullptr))
        return 3;
    FILE *g = fopen(argv[2], "wb");
    fwrite(heff.data(), 4, n, g);
    fwrite(dn.data(), 4, n, g);
    fwrite(dh.data(), 4, n, g);
    fwrite(grain.data(), 4, n, g);
    fwrite(out.data(), 1, n, g);
    fwrite(fs.data(), 4, 4, g);
    fclose(g);
    return 0;
}
