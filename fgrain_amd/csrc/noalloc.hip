// fg_expose_noalloc: the transport of one tile without any device allocation, for
// callers that cannot allocate (code under HIP graph capture, serving loops that
// forbid hipMalloc).  Inputs and outputs are the caller's device memory; the only
// memory beyond them is a fixed 64-byte block per GPU (the persistent kernel's work
// counter and its statistics counters), a static device array whose address is
// looked up once per engine at creation.  Every call on a GPU shares that block, so
// calls that overlap on one GPU race; each call clears it with a fill kernel first
// (no memset node, so the sequence can be captured and replayed).
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>
#include <string>

#include "engine_state.h"

using namespace fg;

namespace {

// statistics counters 0..5 (bounces, nodes, crystal tests, absorbing hits, latent
// photons, global atomics), the persistent kernel's work counter at kWorkWord, packets
// started at 7
constexpr uint32_t kScratchWords = 8;
constexpr uint32_t kWorkWord = 6;
static_assert(kScratchWords * sizeof(unsigned long long) <= FG_NOALLOC_SCRATCH_BYTES,
              "scratch block larger than the documented maximum");

// Host code never uses this name as an address: fg_engine::noalloc_scratch holds the
// device address of the current GPU's copy (noalloc_init).
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

__global__ void __launch_bounds__(64) clear_words_kernel(unsigned long long *p, uint32_t n)
{
    const uint32_t i = threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

// Called by fg_engine_create with the engine's device current.
fg_status noalloc_init(fg_engine *E)
{
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(FG_ERR_DEVICE_REQUEST,
                    std::string("scratch block address: ") + hipGetErrorString(e));
    E->noalloc_scratch = p;
    return FG_OK;
}

extern "C" fg_status fg_expose_noalloc(fg_engine *E, const fg_emulsion *em_h, const fg_device_tile *t,
                                       const fg_exposure_config *cfg, uint64_t seed, void *stream,
                                       fg_expose_stats *stats)
{
    std::string msg;
    if (!E || !em_h || !t) return fail(FG_ERR_INVALID_ARG, "fg_expose_noalloc: null engine, emulsion or tile");
    if (!valid_exposure(cfg, msg)) return fail(FG_ERR_INVALID_ARG, "fg_expose_noalloc: " + msg);
    const Emulsion &em = em_h->em;
    if ((uint64_t)t->n_nodes != (uint64_t)em.nodes.size() || (int64_t)t->n_crystals != em.n)
        return fail(FG_ERR_INVALID_ARG,
                    "fg_expose_noalloc: tile has " + std::to_string(t->n_nodes) + " nodes and " +
                        std::to_string(t->n_crystals) + " crystals, the emulsion " +
                        std::to_string(em.nodes.size()) + " and " + std::to_string(em.n));
    if (t->n_crystals > 0 && (!t->nodes || !t->records || !t->counts))
        return fail(FG_ERR_INVALID_ARG, "fg_expose_noalloc: null node, record or count array");
    if (!offsets_fit_32(t->n_nodes, t->n_crystals))
        return fail(FG_ERR_INVALID_ARG, "fg_expose_noalloc: node or crystal arrays of 4 GiB or more");
    if (t->exposure && (t->tw == 0 || t->th == 0))
        return fail(FG_ERR_INVALID_ARG, "fg_expose_noalloc: empty exposure map");
    if (cfg->photons_per_pixel > 0 && !t->exposure)
        return fail(FG_ERR_INVALID_ARG, "fg_expose_noalloc: photons_per_pixel needs an exposure map");
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    if (!E->noalloc_scratch) return fail(FG_ERR_DEVICE_REQUEST, "fg_expose_noalloc: no scratch block");
    unsigned long long *sc = static_cast<unsigned long long *>(E->noalloc_scratch);
    const hipStream_t s = stream ? static_cast<hipStream_t>(stream) : E->stream;

    TransportArgs a{};
    a.nodes = static_cast<const Node *>(t->nodes);
    a.recs = static_cast<const float4 *>(t->records);
    a.n_nodes = t->n_nodes;
    fill_transport(a, em.cfg, *cfg, t->exposure ? t->tw : 1, t->exposure ? t->th : 1, seed, em.reach,
                   E->tuning);
    a.emap = t->exposure;
    a.estride = t->exposure ? t->tw : 1;
    a.counts = t->counts;
    a.stats = sc;

    (void)hipGetLastError();  // the check below reports this launch only
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, s, sc, kScratchWords);
    HIP_TRY(hipGetLastError(), FG_ERR_PIPELINE);
    if (stats) HIP_TRY(hipEventRecord(E->ev0, s), FG_ERR_CHANNEL);
    if (E->tuning.transport_kernel == 1)
        HIP_TRY(launch_transport(a, s), FG_ERR_PIPELINE);
    else
        HIP_TRY(launch_transport_persistent(a, sc + kWorkWord, E->persistent_grid, s), FG_ERR_PIPELINE);
    if (!stats) return FG_OK;
    HIP_TRY(hipEventRecord(E->ev1, s), FG_ERR_CHANNEL);
    unsigned long long st[kScratchWords];
    HIP_TRY(hipMemcpyAsync(st, sc, sizeof st, hipMemcpyDeviceToHost, s), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(s), FG_ERR_CHANNEL);
    memset(stats, 0, sizeof *stats);
    stats->launched = a.n_photons;
    stats->bounces = st[0];
    stats->nodes_visited = st[1];
    stats->prims_tested = st[2];
    stats->abs_events = st[3];
    stats->abs_total = st[4];
    stats->atomics = st[5];
    stats->packets_run = st[7];
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, E->ev0, E->ev1);
    stats->kernel_ms = ms;
    return FG_OK;
}
