// Engine and emulsion state shared by the C-ABI translation units (engine.cpp,
// adjoint_host.cpp); not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "film.h"
#include "internal.h"
#include "kernels.h"

// Sets the thread-local message fg_last_error returns and passes the status through.
fg_status fail(fg_status s, const std::string &msg);

#define HIP_TRY(call, status)                                                          \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(status, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        const size_t b = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) cap = b;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// The transport addresses nodes (32 B) and crystal records (48 B) by 32-bit byte
// offsets from the start of their arrays (raw buffer loads).
inline bool offsets_fit_32(uint64_t n_nodes, uint64_t n_records)
{
    return n_nodes * 32ull < (1ull << 32) && n_records * 48ull < (1ull << 32);
}

namespace fg {
struct DeviceEmulsion {
    DevBuf nodes, recs, counts;
    DevBuf pre;   // the transport's pre-test records (PreRec), for the replay kernels
    uint32_t n_nodes = 0;
    int64_t n = 0;
    // expected-model scan BVH: every crystal, boxes for the largest boosted expansion,
    // records with semi-axes times exp_fe and sens words holding transport slots; touch
    // graph (CSR, slot order) for the expected clump boost.  Built on first use for a
    // given (expansion, proximity factor).
    DevBuf exp_nodes, exp_recs, exp_ptr, exp_idx;
    uint32_t exp_n_nodes = 0;
    double exp_fe = -1.0, exp_pf = -1.0;
};
}  // namespace fg

struct fg_engine {
    int device = 0;
    int cus = 0;
    int persistent_grid = 0;
    int replay_grid = 0;       // resident blocks of the persistent replay kernels
    fg_tuning tuning{0u, 64u, 80u, 0u, 6u, 0u};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    std::string name;
    std::mutex mu;
    DevBuf emap, stats, scan_nodes, scan_qnodes, scan_recs, scan_consts, scan_lo, out, out2, hp, rays_o, rays_d, hits_n, hits_i,
        hits_t0, hits_t1;
    // expected model and its adjoint
    DevBuf dose64, dose32, prob, dprob, gprob, gdose, gmap, partials, pgrad, gout, boost, dboost,
        gboost;
    DevBuf synth_cells;        // device synthesis grid
    // device address of this GPU's scratch block of fg_expose_noalloc (noalloc.hip)
    void *noalloc_scratch = nullptr;
    // full-frame film compositing; film_mu (taken before mu) serialises the film calls,
    // which hold these buffers across several engine calls
    std::mutex film_mu;
    DevBuf film_rgb, film_gray, film_heff, film_bbox, film_tiles, film_dnorm, film_tmp, film_dhigh,
        film_partials, film_fstats, film_kern, film_layout, film_grain, film_out;
};

struct fg_emulsion {
    fg::Emulsion em;
};

fg_status ensure_device(fg_engine *E, fg::Emulsion &em);
// Looks up the current device's scratch block of fg_expose_noalloc (noalloc.hip); called
// once per engine by fg_engine_create with the engine's device current.
fg_status noalloc_init(fg_engine *E);
namespace fg {
// Tables of an overlapped layout (film_host.cpp) uploaded into buf on stream (complete
// on return); L points into buf.
hipError_t film_layout_upload(uint32_t W, uint32_t H, uint32_t step, uint32_t pad, DevBuf &buf,
                              FilmLayout &L, hipStream_t stream);
}  // namespace fg
// EmulsionConfig of a tile render (render_tile_cpu_full's preset, image_pipeline.rs:925-938)
fg_emulsion_config tile_emulsion_config(const fg_tile_params &p);
// the frame's tile D_neg buffer and its layout (null for abutting tiles)
const float *frame_tile_dneg(const fg_frame *F);
const fg::FilmLayout *frame_layout(const fg_frame *F);
void fill_transport(fg::TransportArgs &a, const fg_emulsion_config &ec, const fg_exposure_config &x,
                    uint32_t tw, uint32_t th, uint64_t seed, double reach, const fg_tuning &tn);
bool valid_exposure(const fg_exposure_config *c, std::string &msg);
bool valid_scan(const fg_scan_config *c, std::string &msg);
