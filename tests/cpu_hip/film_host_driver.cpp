// Drives the host build of film_host.cpp and film.hip (against the HIP shim) through the
// library's own entry points, fg_film_optics and fg_film_composite, on one engine: the
// jobs of the input file run in order, as a test process would call them, so the
// engine's buffers grow and are reused as they are on the device.  Input: jobs of
// (u32 kind 0 optics / 1 composite, u32 W, u32 H, fg_film_options, u8 gray[W H],
// then for a composite f32 h_eff[W H], i64 n_tiles, f32 tiles[n_tiles]).  Output per job:
// optics: f32 h_eff[W H]; composite: u8 out[W H], f32 d_norm, d_high, grain [W H] each,
// f32 high_mean, high_std, norm_factor.  Exit status: 0, or the failing job's status.
#include <cstdio>
#include <string>
#include <vector>

#include "engine_state.h"

static thread_local std::string g_err;
fg_status fail(fg_status s, const std::string &msg)
{
    g_err = msg;
    return s;
}

// Frame and tile entry points fg_render_film_bw calls; the harness drives the stages
// directly, so these are never reached.
extern "C" fg_status fg_tile_params_init(fg_tile_params *, const char *, double, uint32_t, uint32_t, double, double,
                                         uint64_t)
{
    return FG_ERR_PIPELINE;
}
fg_emulsion_config tile_emulsion_config(const fg_tile_params &) { return fg_emulsion_config{}; }
extern "C" fg_status fg_frame_create(fg_engine *, const fg_frame_desc *, fg_frame **) { return FG_ERR_PIPELINE; }
extern "C" fg_status fg_frame_render(fg_engine *, fg_frame *, uint32_t, fg_frame_stats *) { return FG_ERR_PIPELINE; }
extern "C" void fg_frame_destroy(fg_frame *) {}
const fg::FilmLayout *frame_layout(const fg_frame *) { return nullptr; }
const float *frame_tile_dneg(const fg_frame *) { return nullptr; }

template <class T> static bool rd(FILE *f, T *p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    FILE *f = fopen(argv[1], "rb"), *g = fopen(argv[2], "wb");
    if (!f || !g) return 2;
    fg_engine E;
    int rc = 0;
    uint32_t hdr[3];
    while (rc == 0 && rd(f, hdr, 3)) {
        const uint32_t kind = hdr[0], W = hdr[1], H = hdr[2];
        const size_t n = (size_t)W * H;
        fg_film_options o;
        std::vector<uint8_t> gray(n);
        if (!rd(f, &o, 1) || !rd(f, gray.data(), n)) return 2;
        if (kind == 0) {
            std::vector<float> heff(n);
            rc = fg_film_optics(&E, gray.data(), W, H, heff.data(), nullptr);
            fwrite(heff.data(), 4, n, g);
        } else {
            std::vector<float> heff(n);
            int64_t nt = 0;
            if (!rd(f, heff.data(), n) || !rd(f, &nt, 1)) return 2;
            std::vector<float> tiles((size_t)nt);
            if (!rd(f, tiles.data(), (size_t)nt)) return 2;
            std::vector<uint8_t> out(n);
            std::vector<float> dn(n), dh(n), gr(n);
            fg_composite_stats st{};
            rc = fg_film_composite(&E, &o, W, H, gray.data(), heff.data(), tiles.data(), nt, out.data(), dn.data(),
                                   dh.data(), gr.data(), &st);
            fwrite(out.data(), 1, n, g);
            fwrite(dn.data(), 4, n, g);
            fwrite(dh.data(), 4, n, g);
            fwrite(gr.data(), 4, n, g);
            const float fs[3] = {st.high_mean, st.high_std, st.norm_factor};
            fwrite(fs, 4, 3, g);
        }
        if (rc) fprintf(stderr, "job %u x %u kind %u: status %d: %s\n", W, H, kind, rc, g_err.c_str());
    }
    for (DevBuf *b : {&E.film_rgb, &E.film_gray, &E.film_heff, &E.film_bbox, &E.film_tiles, &E.film_dnorm,
                      &E.film_tmp, &E.film_dhigh, &E.film_partials, &E.film_fstats, &E.film_kern, &E.film_layout,
                      &E.film_grain, &E.film_out})
        b->release();
    fclose(f);
    fclose(g);
    return rc;
}
