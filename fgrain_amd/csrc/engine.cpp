// C ABI of the MI355X engine (include/fgrain_hip.h): engine lifecycle, the library
// stages, the drop-in tile render and device-resident frames.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batched.h"
#include "engine_state.h"
#include "synth.h"

using namespace fg;

// ------------------------------------------------------------------ errors -------
static thread_local std::string g_last_error;

fg_status fail(fg_status s, const std::string &msg)
{
    g_last_error = msg;
    return s;
}

// ------------------------------------------------------------------ engine -------
extern "C" uint32_t fg_abi_version(void) { return FG_ABI_VERSION; }

extern "C" const char *fg_last_error(void) { return g_last_error.c_str(); }

extern "C" fg_status fg_engine_create(fg_engine **out)
{
    if (!out) return fail(FG_ERR_INVALID_ARG, "fg_engine_create: null output");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(FG_ERR_ADAPTER_NOT_FOUND,
                    std::string("no HIP device: ") + (e != hipSuccess ? hipGetErrorString(e) : "0 devices"));
    fg_engine *E = new fg_engine();
    e = hipGetDevice(&E->device);
    if (e == hipSuccess) e = hipSetDevice(E->device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, E->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&E->ev0);
    if (e == hipSuccess) e = hipEventCreate(&E->ev1);
    if (e == hipSuccess) e = hipEventCreate(&E->ev2);
    if (e == hipSuccess) e = hipEventCreate(&E->ev3);
    if (e != hipSuccess) {
        delete E;
        return fail(FG_ERR_DEVICE_REQUEST, std::string("HIP device setup: ") + hipGetErrorString(e));
    }
    E->name = std::string(prop.name) + " (" + prop.gcnArchName + ", " +
              std::to_string(prop.multiProcessorCount) + " CUs)";
    E->cus = prop.multiProcessorCount;
    E->persistent_grid = E->cus * transport_persistent_blocks_per_cu();
    if (noalloc_init(E) != FG_OK) {
        const std::string m = g_last_error;
        fg_engine_destroy(E);
        return fail(FG_ERR_DEVICE_REQUEST, m);
    }
    *out = E;
    return FG_OK;
}

extern "C" void fg_engine_destroy(fg_engine *E)
{
    if (!E) return;
    for (DevBuf *b : {&E->emap, &E->stats, &E->scan_nodes, &E->scan_qnodes, &E->scan_recs, &E->scan_consts, &E->scan_lo, &E->out, &E->out2, &E->hp,
                      &E->rays_o, &E->rays_d, &E->hits_n, &E->hits_i, &E->hits_t0, &E->hits_t1,
                      &E->dose64, &E->dose32, &E->prob, &E->dprob, &E->gprob, &E->gdose,
                      &E->gmap, &E->partials, &E->pgrad, &E->gout, &E->boost, &E->dboost,
                      &E->gboost, &E->synth_cells, &E->film_rgb, &E->film_gray, &E->film_heff, &E->film_bbox,
                      &E->film_tiles, &E->film_dnorm, &E->film_tmp, &E->film_dhigh, &E->film_partials,
                      &E->film_fstats, &E->film_kern, &E->film_layout, &E->film_grain, &E->film_out})
        b->release();
    if (E->ev0) (void)hipEventDestroy(E->ev0);
    if (E->ev1) (void)hipEventDestroy(E->ev1);
    if (E->ev2) (void)hipEventDestroy(E->ev2);
    if (E->ev3) (void)hipEventDestroy(E->ev3);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
}

extern "C" const char *fg_engine_adapter_name(const fg_engine *E) { return E ? E->name.c_str() : ""; }

extern "C" fg_status fg_engine_set_tuning(fg_engine *E, const fg_tuning *t)
{
    if (!E || !t) return fail(FG_ERR_INVALID_ARG, "fg_engine_set_tuning: null argument");
    if (t->transport_kernel > 1 || t->transport_wait_pct > 100 || t->scan_rays_per_thread > 64 ||
        (t->transport_waves != 0 && t->transport_waves != 6 && t->transport_waves != 7) ||
        t->transport_prereduce > 1)
        return fail(FG_ERR_INVALID_ARG,
                    "fg_engine_set_tuning: transport_kernel must be 0 or 1, wait_pct <= 100, "
                    "scan_rays_per_thread <= 64, transport_waves 0, 6 or 7, transport_prereduce 0 or 1");
    std::lock_guard<std::mutex> lk(E->mu);
    E->tuning.transport_kernel = t->transport_kernel;
    E->tuning.transport_chunk = t->transport_chunk ? std::max(t->transport_chunk, 64u) : 64u;
    E->tuning.transport_wait_pct = t->transport_wait_pct ? t->transport_wait_pct : 80u;
    E->tuning.scan_rays_per_thread = t->scan_rays_per_thread;
    E->tuning.transport_waves = t->transport_waves ? t->transport_waves : 6u;
    E->tuning.transport_prereduce = t->transport_prereduce;
    return FG_OK;
}

extern "C" fg_status fg_engine_get_tuning(const fg_engine *E, fg_tuning *t)
{
    if (!E || !t) return fail(FG_ERR_INVALID_ARG, "fg_engine_get_tuning: null argument");
    *t = E->tuning;
    return FG_OK;
}

static std::once_flag g_engine_once;
static fg_engine *g_engine = nullptr;
static fg_status g_engine_status = FG_OK;
static std::string g_engine_msg;

extern "C" fg_status fg_get_engine(fg_engine **out)
{
    if (!out) return fail(FG_ERR_INVALID_ARG, "fg_get_engine: null output");
    std::call_once(g_engine_once, [] {
        g_engine_status = fg_engine_create(&g_engine);
        if (g_engine_status != FG_OK) g_engine_msg = g_last_error;
    });
    *out = g_engine;
    if (g_engine_status != FG_OK) return fail(g_engine_status, g_engine_msg);
    return FG_OK;
}

extern "C" fg_status fg_calibrate_valu(fg_engine *E, uint32_t blocks, uint32_t iters, float *ms_kernel,
                                      uint64_t *valu_per_wave)
{
    if (!E || blocks == 0 || iters == 0 || blocks > (1u << 20))
        return fail(FG_ERR_INVALID_ARG, "fg_calibrate_valu: engine, 1..2^20 blocks and iters > 0 required");
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    HIP_TRY(E->out.ensure((size_t)blocks * sizeof(float)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipEventRecord(E->ev0, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(launch_valu_probe(E->out.as<float>(), blocks, iters, E->stream), FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev1, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, E->ev0, E->ev1), FG_ERR_CHANNEL);
    if (ms_kernel) *ms_kernel = ms;
    if (valu_per_wave) *valu_per_wave = 64ull * iters + 4ull;
    return FG_OK;
}

// ---------------------------------------------------------------- emulsion -------
static bool valid_cfg(const fg_emulsion_config *c, std::string &msg)
{
    if (!c) { msg = "null emulsion config"; return false; }
    if (!(c->width >= 0 && c->height >= 0 && c->depth > 0)) { msg = "emulsion extents must be >= 0 (depth > 0)"; return false; }
    if (!(c->r_min > 0)) { msg = "r_min must be > 0"; return false; }
    if (c->morphology > 2) { msg = "morphology must be 0, 1 or 2"; return false; }
    if (c->morphology == 1 && !(c->tabular_aspect_ratio > 0)) { msg = "tabular aspect ratio must be > 0"; return false; }
    const double cells = (c->width / (c->r_min / sqrt(3.0)) + 1) * (c->height / (c->r_min / sqrt(3.0)) + 1) *
                         (c->depth / (c->r_min / sqrt(3.0)) + 1);
    if (cells > 4e9) { msg = "emulsion too large for one Poisson grid"; return false; }
    return true;
}

extern "C" fg_status fg_emulsion_synthesize(const fg_emulsion_config *cfg, uint64_t seed,
                                            fg_emulsion **out)
{
    std::string msg;
    if (!out || !valid_cfg(cfg, msg)) return fail(FG_ERR_INVALID_ARG, "fg_emulsion_synthesize: " + msg);
    fg_emulsion *E = new fg_emulsion();
    synthesize(*cfg, seed, E->em);
    if (!finalize_emulsion(E->em, msg)) {
        delete E;
        return fail(FG_ERR_INVALID_ARG, "fg_emulsion_synthesize: " + msg);
    }
    *out = E;
    return FG_OK;
}

// Dart-throwing grid of one emulsion (synth.hip): cell r / sqrt(3), extents in float.
static bool dart_grid(const fg_emulsion_config &cfg, uint64_t seed, DartArgs &a)
{
    a = DartArgs{};
    if (!(cfg.width > 0 && cfg.height > 0 && cfg.depth > 0)) return false;
    const float cellf = (float)(cfg.r_min / sqrt(3.0));
    a.gx = (uint32_t)ceil(cfg.width / (double)cellf);
    a.gy = (uint32_t)ceil(cfg.height / (double)cellf);
    a.gz = (uint32_t)ceil(cfg.depth / (double)cellf);
    a.cell = cellf;
    a.r2 = (float)(cfg.r_min * cfg.r_min);
    a.W = (float)cfg.width;
    a.H = (float)cfg.height;
    a.D = (float)cfg.depth;
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    return true;
}

static void compact_cells(const float4 *cells, size_t n, std::vector<double> &pts)
{
    pts.clear();
    for (size_t c = 0; c < n; c++)
        if (cells[c].w != 0.0f) {
            pts.push_back(cells[c].x);
            pts.push_back(cells[c].y);
            pts.push_back(cells[c].z);
        }
}

// Poisson-disk centres of many emulsions in one pass of device launches into
// E->synth_cells (tile k's grid at offsets[k]); with cells given, they come back to the
// host in one buffer.  Caller holds E->mu.
static fg_status device_cells_tiles(fg_engine *E, const std::vector<fg_emulsion_config> &cfgs,
                                    const std::vector<uint64_t> &seeds, std::vector<float4> *cells,
                                    std::vector<size_t> &offsets)
{
    const size_t n = cfgs.size();
    std::vector<DartArgs> args(n);
    offsets.assign(n + 1, 0);
    uint32_t max_phase = 0;
    for (size_t k = 0; k < n; k++) {
        DartArgs &a = args[k];
        size_t cells_k = 0;
        if (dart_grid(cfgs[k], seeds[k], a)) {
            const double g = (double)a.gx * a.gy * a.gz;
            if (g > (double)(1u << 28))
                return fail(FG_ERR_INVALID_ARG, "device synthesis: tile grid too large");
            cells_k = (size_t)a.gx * a.gy * a.gz;
            max_phase = std::max(max_phase, ((a.gx + 2u) / 3u) * ((a.gy + 2u) / 3u) * ((a.gz + 2u) / 3u));
        }
        offsets[k + 1] = offsets[k] + cells_k;
    }
    const size_t total = offsets[n];
    HIP_TRY(E->synth_cells.ensure(std::max<size_t>(total, 1) * sizeof(float4)), FG_ERR_BUFFER_RANGE);
    for (size_t k = 0; k < n; k++) args[k].cells = E->synth_cells.as<float4>() + offsets[k];
    DevBuf dargs;
    HIP_TRY(dargs.ensure(n * sizeof(DartArgs)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemcpyAsync(dargs.p, args.data(), n * sizeof(DartArgs), hipMemcpyHostToDevice, E->stream),
            FG_ERR_BUFFER_MAP);
    if (total) HIP_TRY(hipMemsetAsync(E->synth_cells.p, 0, total * sizeof(float4), E->stream), FG_ERR_BUFFER_MAP);
    const uint32_t rounds = (n && cfgs[0].k_samples) ? cfgs[0].k_samples : 30u;
    fg_status st = FG_OK;
    hipError_t e = launch_darts_tiles(dargs.as<DartArgs>(), (uint32_t)n, max_phase, rounds, E->stream);
    if (cells) {
        cells->resize(total);
        if (e == hipSuccess && total)
            e = hipMemcpyAsync(cells->data(), E->synth_cells.p, total * sizeof(float4), hipMemcpyDeviceToHost,
                               E->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(E->stream);
    if (e != hipSuccess) st = fail(FG_ERR_PIPELINE, std::string("device synthesis: ") + hipGetErrorString(e));
    dargs.release();
    return st;
}

// Poisson-disk centres on the device (synth.hip), in cell order.
static fg_status device_points(fg_engine *E, const fg_emulsion_config &cfg, uint64_t seed,
                               std::vector<double> &pts)
{
    std::vector<float4> cells;
    std::vector<size_t> off;
    fg_status s = device_cells_tiles(E, {cfg}, {seed}, &cells, off);
    if (s != FG_OK) return s;
    compact_cells(cells.data(), cells.size(), pts);
    return FG_OK;
}

extern "C" fg_status fg_emulsion_synthesize_device(fg_engine *E, const fg_emulsion_config *cfg,
                                                   uint64_t seed, fg_emulsion **out)
{
    std::string msg;
    if (!E || !out || !valid_cfg(cfg, msg))
        return fail(FG_ERR_INVALID_ARG, "fg_emulsion_synthesize_device: " + (msg.empty() ? std::string("null argument") : msg));
    *out = nullptr;
    std::vector<double> pts;
    {
        std::lock_guard<std::mutex> lk(E->mu);
        HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
        fg_status s = device_points(E, *cfg, seed, pts);
        if (s != FG_OK) return s;
    }
    fg_emulsion *em = new fg_emulsion();
    em->em.cfg = *cfg;
    em->em.n = (int64_t)(pts.size() / 3);
    em->em.pos = std::move(pts);
    attributes_per_point(*cfg, seed, em->em);
    if (!finalize_emulsion(em->em, msg)) {
        delete em;
        return fail(FG_ERR_INVALID_ARG, "fg_emulsion_synthesize_device: " + msg);
    }
    *out = em;
    return FG_OK;
}

extern "C" fg_status fg_emulsion_from_crystals(const fg_emulsion_config *cfg, int64_t n,
                                               const double *pos, const double *dims,
                                               const double *quat, const int32_t *morph,
                                               const double *sens, fg_emulsion **out)
{
    std::string msg;
    if (!out || n < 0 || (n > 0 && (!pos || !dims || !quat || !morph || !sens)) || !cfg)
        return fail(FG_ERR_INVALID_ARG, "fg_emulsion_from_crystals: null or negative argument");
    if (n > (int64_t)(1u << 27)) return fail(FG_ERR_INVALID_ARG, "fg_emulsion_from_crystals: too many crystals");
    for (int64_t i = 0; i < n; i++)
        if (morph[i] < 0 || morph[i] > 2) return fail(FG_ERR_INVALID_ARG, "fg_emulsion_from_crystals: bad morphology");
    fg_emulsion *E = new fg_emulsion();
    Emulsion &em = E->em;
    em.cfg = *cfg;
    em.n = n;
    em.pos.assign(pos, pos + 3 * n);
    em.dims.assign(dims, dims + 3 * n);
    em.quat.assign(quat, quat + 4 * n);
    em.morph.assign(morph, morph + n);
    em.sens.assign(sens, sens + n);
    if (!finalize_emulsion(em, msg)) {
        delete E;
        return fail(FG_ERR_INVALID_ARG, "fg_emulsion_from_crystals: " + msg);
    }
    *out = E;
    return FG_OK;
}

static void free_device(Emulsion &em)
{
    if (em.dev) {
        em.dev->nodes.release();
        em.dev->recs.release();
        em.dev->pre.release();
        em.dev->counts.release();
        em.dev->exp_nodes.release();
        em.dev->exp_recs.release();
        em.dev->exp_ptr.release();
        em.dev->exp_idx.release();
        delete em.dev;
        em.dev = nullptr;
    }
}

extern "C" void fg_emulsion_destroy(fg_emulsion *E)
{
    if (!E) return;
    free_device(E->em);
    delete E;
}

extern "C" int64_t fg_emulsion_count(const fg_emulsion *E) { return E ? E->em.n : 0; }

extern "C" fg_status fg_emulsion_get_crystals(const fg_emulsion *E, double *pos, double *dims,
                                              double *quat, int32_t *morph, double *sens)
{
    if (!E) return fail(FG_ERR_INVALID_ARG, "fg_emulsion_get_crystals: null emulsion");
    const Emulsion &em = E->em;
    if (pos) memcpy(pos, em.pos.data(), em.pos.size() * sizeof(double));
    if (dims) memcpy(dims, em.dims.data(), em.dims.size() * sizeof(double));
    if (quat) memcpy(quat, em.quat.data(), em.quat.size() * sizeof(double));
    if (morph) memcpy(morph, em.morph.data(), em.morph.size() * sizeof(int32_t));
    if (sens) memcpy(sens, em.sens.data(), em.sens.size() * sizeof(double));
    return FG_OK;
}

extern "C" fg_status fg_emulsion_get_bvh(const fg_emulsion *E, int64_t *n_nodes, double *bmin,
                                         double *bmax, uint32_t *left, uint32_t *right,
                                         uint32_t *offset, uint32_t *count, uint32_t *indices)
{
    if (!E || !n_nodes) return fail(FG_ERR_INVALID_ARG, "fg_emulsion_get_bvh: null argument");
    const RefBvh &b = E->em.bvh;
    const int64_t nn = b.nodes();
    *n_nodes = nn;
    if (bmin) memcpy(bmin, b.mn.data(), (size_t)nn * 3 * sizeof(double));
    if (bmax) memcpy(bmax, b.mx.data(), (size_t)nn * 3 * sizeof(double));
    if (left) memcpy(left, b.left.data(), (size_t)nn * sizeof(uint32_t));
    if (right) memcpy(right, b.right.data(), (size_t)nn * sizeof(uint32_t));
    if (offset) memcpy(offset, b.offset.data(), (size_t)nn * sizeof(uint32_t));
    if (count) memcpy(count, b.count.data(), (size_t)nn * sizeof(uint32_t));
    if (indices) memcpy(indices, b.idx.data(), b.idx.size() * sizeof(uint32_t));
    return FG_OK;
}

extern "C" fg_status fg_emulsion_get_device_layout(const fg_emulsion *E, void *nodes, void *records)
{
    if (!E) return fail(FG_ERR_INVALID_ARG, "fg_emulsion_get_device_layout: null emulsion");
    if (nodes) memcpy(nodes, E->em.nodes.data(), E->em.nodes.size() * sizeof(Node));
    if (records) memcpy(records, E->em.recs.data(), E->em.recs.size() * sizeof(Crystal));
    return FG_OK;
}

extern "C" fg_status fg_emulsion_get_state(const fg_emulsion *E, uint32_t *counts,
                                           uint8_t *developed, double *fe, int64_t *clump_id)
{
    if (!E) return fail(FG_ERR_INVALID_ARG, "fg_emulsion_get_state: null emulsion");
    const Emulsion &em = E->em;
    if (counts) memcpy(counts, em.counts.data(), em.counts.size() * sizeof(uint32_t));
    if (developed) memcpy(developed, em.developed.data(), em.developed.size());
    if (fe) memcpy(fe, em.fe.data(), em.fe.size() * sizeof(double));
    if (clump_id) memcpy(clump_id, em.clump.data(), em.clump.size() * sizeof(int64_t));
    return FG_OK;
}

extern "C" fg_status fg_emulsion_set_state(fg_emulsion *E, const uint32_t *counts,
                                           const uint8_t *developed, const double *fe)
{
    if (!E) return fail(FG_ERR_INVALID_ARG, "fg_emulsion_set_state: null emulsion");
    Emulsion &em = E->em;
    if (counts) em.counts.assign(counts, counts + em.n);
    if (developed) em.developed.assign(developed, developed + em.n);
    if (fe) em.fe.assign(fe, fe + em.n);
    em.dev_state_dirty = true;
    return FG_OK;
}

// ------------------------------------------------------------- device helpers ---
fg_status ensure_device(fg_engine *E, Emulsion &em)
{
    if (em.dev) return FG_OK;
    if (!offsets_fit_32(em.nodes.size(), em.recs.size()))
        return fail(FG_ERR_INVALID_ARG, "emulsion: node or crystal arrays of 4 GiB or more");
    em.dev = new DeviceEmulsion();
    DeviceEmulsion &d = *em.dev;
    d.n = em.n;
    d.n_nodes = (uint32_t)em.nodes.size();
    HIP_TRY(d.nodes.ensure(em.nodes.size() * sizeof(Node)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(d.recs.ensure(em.recs.size() * sizeof(Crystal)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(d.counts.ensure((size_t)em.n * sizeof(uint32_t)), FG_ERR_BUFFER_RANGE);
    if (!em.nodes.empty())
        HIP_TRY(hipMemcpyAsync(d.nodes.p, em.nodes.data(), em.nodes.size() * sizeof(Node),
                               hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    if (!em.recs.empty())
        HIP_TRY(hipMemcpyAsync(d.recs.p, em.recs.data(), em.recs.size() * sizeof(Crystal),
                               hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(d.pre.ensure(std::max<size_t>(em.recs.size(), 1) * sizeof(PreRec)), FG_ERR_BUFFER_RANGE);
    if (!em.recs.empty()) {
        std::vector<PreRec> pre(em.recs.size());
        for (size_t k = 0; k < pre.size(); k++) pre[k] = pre_record(em.recs[k].c, em.recs[k].h, em.recs[k].morph);
        HIP_TRY(hipMemcpyAsync(d.pre.p, pre.data(), pre.size() * sizeof(PreRec), hipMemcpyHostToDevice, E->stream),
                FG_ERR_BUFFER_MAP);
        HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);   // before pre goes out of scope
    }
    return FG_OK;
}

static bool g_full_traversal = getenv("FG_TRANSPORT_FULL_TRAVERSAL") != nullptr;

void fill_transport(TransportArgs &a, const fg_emulsion_config &ec, const fg_exposure_config &x,
                    uint32_t tw, uint32_t th, uint64_t seed, double reach, const fg_tuning &tn)
{
    a.chunk = tn.transport_chunk;
    a.wait_pct = tn.transport_wait_pct;
    a.prereduce = tn.transport_prereduce;
    a.base = (float)x.base_thickness;
    a.reach = g_full_traversal ? -1.0f : (float)reach * 1.0001f;
    const fg_optics &o = ec.optics;
    const double mu_t = o.absorption_coeff + o.scattering_coeff;
    const double r = (1.0 - o.refractive_index) / (1.0 + o.refractive_index);
    a.W = (float)ec.width;
    a.H = (float)ec.height;
    a.D = (float)ec.depth;
    a.pw = (float)(ec.width / (double)std::max(tw, 1u));
    a.ph = (float)(ec.height / (double)std::max(th, 1u));
    a.mu_t = (float)mu_t;
    a.albedo = (float)(mu_t > 0.0 ? o.scattering_coeff / mu_t : 0.0);
    a.r0 = (float)(r * r);
    a.hal_eff = (float)(o.antihalation_reflectance * pow(10.0, -2.0 * o.antihalation_density));
    a.mu_abs = (float)x.crystal_absorption_coeff;
    a.theta_max = (float)x.illumination_angle_rad;
    a.g = (float)o.anisotropy_g;
    a.step_cap = (float)(ec.depth * 1.5);
    a.max_bounces = x.max_bounces;
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    a.tw = std::max(tw, 1u);
    a.th = std::max(th, 1u);
    a.ppp = x.photons_per_pixel;
    if (x.photons_per_pixel > 0)
        a.n_photons = (uint64_t)x.photons_per_pixel * a.tw * a.th;
    else
        a.n_photons = (uint64_t)llround(ec.width * ec.height * x.photons_per_sq_unit);
}

// The transport kernel the tuning selects; work is a zeroed 64-bit counter.
static hipError_t launch_transport_any(fg_engine *E, const TransportArgs &a, unsigned long long *work)
{
    if (E->tuning.transport_kernel == 1) return launch_transport(a, E->stream);
    return launch_transport_persistent(a, work, E->persistent_grid, E->stream);
}

static fg_status expose_locked(fg_engine *E, Emulsion &em, const float *emap, uint32_t tw,
                               uint32_t th, const fg_exposure_config &cfg, uint64_t seed,
                               fg_expose_stats *stats)
{
    fg_status s = ensure_device(E, em);
    if (s != FG_OK) return s;
    DeviceEmulsion &d = *em.dev;
    TransportArgs a{};
    a.nodes = d.nodes.as<Node>();
    a.recs = d.recs.as<float4>();
    a.n_nodes = d.n_nodes;
    fill_transport(a, em.cfg, cfg, emap ? tw : 1, emap ? th : 1, seed, em.reach, E->tuning);
    if (emap) {
        HIP_TRY(E->emap.ensure((size_t)tw * th * sizeof(float)), FG_ERR_BUFFER_RANGE);
        HIP_TRY(hipMemcpyAsync(E->emap.p, emap, (size_t)tw * th * sizeof(float), hipMemcpyHostToDevice,
                               E->stream), FG_ERR_BUFFER_MAP);
        a.emap = E->emap.as<float>();
        a.estride = tw;
    }
    HIP_TRY(E->stats.ensure(8 * sizeof(unsigned long long)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemsetAsync(E->stats.p, 0, 8 * sizeof(unsigned long long), E->stream), FG_ERR_BUFFER_MAP);
    a.estride = emap ? tw : 1;
    if (em.n > 0) HIP_TRY(hipMemsetAsync(d.counts.p, 0, (size_t)em.n * 4, E->stream), FG_ERR_BUFFER_MAP);
    a.counts = d.counts.as<uint32_t>();
    a.stats = E->stats.as<unsigned long long>();
    HIP_TRY(hipEventRecord(E->ev0, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(launch_transport_any(E, a, E->stats.as<unsigned long long>() + 6), FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev1, E->stream), FG_ERR_CHANNEL);
    std::vector<uint32_t> c((size_t)em.n);
    unsigned long long st[8];
    if (em.n > 0)
        HIP_TRY(hipMemcpyAsync(c.data(), d.counts.p, (size_t)em.n * 4, hipMemcpyDeviceToHost, E->stream),
                FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpyAsync(st, E->stats.p, sizeof st, hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    for (size_t k = 0; k < c.size(); k++) em.counts[em.bvh.idx[k]] += c[k];
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->launched = a.n_photons;
        stats->bounces = st[0];
        stats->nodes_visited = st[1];
        stats->prims_tested = st[2];
        stats->abs_events = st[3];
        stats->abs_total = st[4];
        stats->atomics = st[5];
        stats->packets_run = st[7];
        float ms = 0;
        (void)hipEventElapsedTime(&ms, E->ev0, E->ev1);
        stats->kernel_ms = ms;
    }
    return FG_OK;
}

bool valid_exposure(const fg_exposure_config *c, std::string &msg)
{
    if (!c) { msg = "null exposure config"; return false; }
    if (!(c->illumination_angle_rad >= 0.0)) { msg = "illumination angle must be >= 0"; return false; }
    if (!(c->photons_per_sq_unit >= 0.0)) { msg = "photon density must be >= 0"; return false; }
    if (!(c->base_thickness >= 0.0)) { msg = "base thickness must be >= 0"; return false; }
    return true;
}

extern "C" fg_status fg_expose(fg_engine *E, fg_emulsion *em, const float *emap, uint32_t tw,
                               uint32_t th, const fg_exposure_config *cfg, uint64_t seed,
                               fg_expose_stats *stats)
{
    std::string msg;
    if (!E || !em) return fail(FG_ERR_INVALID_ARG, "fg_expose: null engine or emulsion");
    if (!valid_exposure(cfg, msg)) return fail(FG_ERR_INVALID_ARG, "fg_expose: " + msg);
    if (emap && (tw == 0 || th == 0)) return fail(FG_ERR_INVALID_ARG, "fg_expose: empty exposure map");
    if (cfg->photons_per_pixel > 0 && !emap)
        return fail(FG_ERR_INVALID_ARG, "fg_expose: photons_per_pixel needs an exposure map");
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    return expose_locked(E, em->em, emap, tw, th, *cfg, seed, stats);
}

static fg_status develop_device(fg_engine *E, Emulsion &em, const fg_development_config &cfg,
                                fg_development_report &rep);

extern "C" fg_status fg_develop(fg_engine *E, fg_emulsion *em, const fg_development_config *cfg,
                                fg_development_report *report)
{
    if (!em || !cfg) return fail(FG_ERR_INVALID_ARG, "fg_develop: null argument");
    fg_development_report rep;
    if (E && em->em.n > 0) {
        const fg_status s = develop_device(E, em->em, *cfg, rep);
        if (s != FG_OK) return s;
    } else {
        develop_host(em->em, *cfg, rep);  // no engine: the host restatement
    }
    if (report) *report = rep;
    return FG_OK;
}

// Developed-crystal BVH (transmittance.rs:82-89) in the threaded layout, its records
// and each record's centre rounding residual (f64 centre minus f32 centre; 0 in w).
struct ScanBvh {
    std::vector<Node> nodes;
    std::vector<Crystal> recs;
    std::vector<float4> lo;
};

static bool build_scan_bvh(const Emulsion &em, ScanBvh &sb, std::string &msg)
{
    std::vector<Node> &nodes = sb.nodes;
    std::vector<Crystal> &recs = sb.recs;
    std::vector<uint32_t> subset;
    for (int64_t i = 0; i < em.n; i++)
        if (em.developed[(size_t)i]) subset.push_back((uint32_t)i);
    RefBvh b;
    build_ref_bvh((int64_t)subset.size(), em.pos.data(), em.dims.data(), em.developed.data(),
                  em.fe.data(), subset.data(), b);
    if (!flatten_bvh(b, nodes, msg)) return false;
    // records in leaf order of the subset BVH
    RefBvh mapped = b;
    for (auto &k : mapped.idx) k = subset[k];
    make_records(em, mapped, em.developed.data(), em.fe.data(), recs);
    sb.lo.resize(recs.size());
    for (size_t k = 0; k < recs.size(); k++) {
        const double *p = &em.pos[3 * (size_t)mapped.idx[k]];
        sb.lo[k] = make_float4((float)(p[0] - (double)recs[k].c[0]), (float)(p[1] - (double)recs[k].c[1]),
                               (float)(p[2] - (double)recs[k].c[2]), 0.0f);
    }
    return true;
}

bool valid_scan(const fg_scan_config *c, std::string &msg)
{
    if (!c) { msg = "null scan config"; return false; }
    if (c->mode > 3) { msg = "scan mode must be 0..3"; return false; }
    if (c->supersample > 8) { msg = "supersample must be <= 8"; return false; }
    if (c->samples_per_pixel > 1024) { msg = "samples_per_pixel must be <= 1024"; return false; }
    if ((uint64_t)c->width_px * c->height_px > (1ull << 31)) { msg = "scan grid too large"; return false; }
    return true;
}

// Enqueues the scan of an emulsion whose developed-crystal BVH (nodes, recs) was built
// on the host by build_scan_bvh, timed by E->ev0 / E->ev1; *res receives the device
// buffer of the result (E->out, or E->out2 after the OTF).  The caller holds E->mu.
static fg_status scan_enqueue(fg_engine *E, const Emulsion &em, const ScanBvh &sb, const fg_scan_config &cfg,
                              uint64_t seed, const float **res_out)
{
    const std::vector<Node> &nodes = sb.nodes;
    const std::vector<Crystal> &recs = sb.recs;
    HIP_TRY(E->scan_nodes.ensure(nodes.size() * sizeof(Node)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->scan_recs.ensure(recs.size() * sizeof(Crystal)), FG_ERR_BUFFER_RANGE);
    if (!nodes.empty())
        HIP_TRY(hipMemcpyAsync(E->scan_nodes.p, nodes.data(), nodes.size() * sizeof(Node),
                               hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    if (!recs.empty())
        HIP_TRY(hipMemcpyAsync(E->scan_recs.p, recs.data(), recs.size() * sizeof(Crystal),
                               hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(E->scan_lo.ensure(recs.size() * sizeof(float4)), FG_ERR_BUFFER_RANGE);
    if (!recs.empty())
        HIP_TRY(hipMemcpyAsync(E->scan_lo.p, sb.lo.data(), recs.size() * sizeof(float4), hipMemcpyHostToDevice,
                               E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(E->scan_consts.ensure(2 * recs.size() * sizeof(float4)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(launch_down_constants(E->scan_recs.as<float4>(), E->scan_lo.as<float4>(), (uint32_t)recs.size(),
                                  E->scan_consts.as<float4>(), E->stream), FG_ERR_PIPELINE);
    const size_t npx = (size_t)cfg.width_px * cfg.height_px;
    HIP_TRY(E->out.ensure(npx * sizeof(float)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->scan_qnodes.ensure(std::max<size_t>(nodes.size(), 1) * sizeof(ScanQNode)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(launch_scan_qnodes(E->scan_nodes.as<Node>(), (uint32_t)nodes.size(), E->scan_qnodes.as<ScanQNode>(),
                               E->stream), FG_ERR_PIPELINE);
    ScanArgs a{};
    a.nodes = E->scan_nodes.as<Node>();
    a.qnodes = E->scan_qnodes.as<ScanQNode>();
    a.recs = E->scan_recs.as<float4>();
    a.consts = E->scan_consts.as<float4>();
    a.n_nodes = (uint32_t)nodes.size();
    a.w_px = cfg.width_px;
    a.h_px = cfg.height_px;
    const uint32_t spp = std::max(cfg.samples_per_pixel, 1u);
    a.g = std::max((uint32_t)llround(sqrt((double)spp)), 1u);
    a.ss = std::max(cfg.supersample, 1u);
    a.mode = cfg.mode;
    a.pw = (float)(em.cfg.width / (double)cfg.width_px);
    a.ph = (float)(em.cfg.height / (double)cfg.height_px);
    a.z0 = (float)(em.cfg.depth + 0.1);
    a.sigma = (float)cfg.silver_opacity;
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    a.out = E->out.as<float>();
    a.ostride = cfg.width_px;
    a.rays_per_lane = E->tuning.scan_rays_per_thread;
    HIP_TRY(hipEventRecord(E->ev0, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(launch_scan(a, E->stream), FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev1, E->stream), FG_ERR_CHANNEL);
    const float *res = E->out.as<float>();
    if (cfg.otf) {
        HIP_TRY(E->out2.ensure(npx * sizeof(float)), FG_ERR_BUFFER_RANGE);
        HIP_TRY(launch_otf3(E->out.as<float>(), E->out2.as<float>(), cfg.width_px, cfg.height_px, E->stream),
                FG_ERR_PIPELINE);
        res = E->out2.as<float>();
    }
    *res_out = res;
    return FG_OK;
}

// scan_enqueue, then the result copied to the host array out (width_px * height_px).
static fg_status scan_locked(fg_engine *E, const Emulsion &em, const ScanBvh &sb, const fg_scan_config &cfg,
                             uint64_t seed, float *out, float *kernel_ms)
{
    const float *res = nullptr;
    fg_status s = scan_enqueue(E, em, sb, cfg, seed, &res);
    if (s != FG_OK) return s;
    const size_t npx = (size_t)cfg.width_px * cfg.height_px;
    HIP_TRY(hipMemcpyAsync(out, res, npx * sizeof(float), hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    if (kernel_ms) (void)hipEventElapsedTime(kernel_ms, E->ev0, E->ev1);
    return FG_OK;
}

extern "C" fg_status fg_scan(fg_engine *E, const fg_emulsion *em, const fg_scan_config *cfg,
                             uint64_t seed, float *out)
{
    std::string msg;
    if (!E || !em || !out) return fail(FG_ERR_INVALID_ARG, "fg_scan: null argument");
    if (!valid_scan(cfg, msg)) return fail(FG_ERR_INVALID_ARG, "fg_scan: " + msg);
    if (cfg->width_px == 0 || cfg->height_px == 0) return FG_OK;
    ScanBvh sb;
    if (!build_scan_bvh(em->em, sb, msg)) return fail(FG_ERR_INVALID_ARG, "fg_scan: " + msg);
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    return scan_locked(E, em->em, sb, *cfg, seed, out, nullptr);
}

static fg_status traverse(fg_engine *E, const fg_emulsion *emw, int64_t n_rays, const float *o,
                          const float *d, int32_t max_hits, int32_t *n_hits, int32_t *idx, float *t0,
                          float *t1, bool first)
{
    if (!E || !emw || n_rays < 0 || (n_rays > 0 && (!o || !d || !idx || !t0 || !t1)) ||
        (!first && (max_hits <= 0 || !n_hits)))
        return fail(FG_ERR_INVALID_ARG, "traverse: bad argument");
    if (n_rays == 0) return FG_OK;
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    Emulsion &em = const_cast<Emulsion &>(emw->em);
    fg_status s = ensure_device(E, em);
    if (s != FG_OK) return s;
    const int64_t per = first ? 1 : max_hits;
    HIP_TRY(E->rays_o.ensure((size_t)n_rays * 12), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->rays_d.ensure((size_t)n_rays * 12), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->hits_n.ensure((size_t)n_rays * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->hits_i.ensure((size_t)(n_rays * per) * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->hits_t0.ensure((size_t)(n_rays * per) * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->hits_t1.ensure((size_t)(n_rays * per) * 4), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemcpyAsync(E->rays_o.p, o, (size_t)n_rays * 12, hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpyAsync(E->rays_d.p, d, (size_t)n_rays * 12, hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    DeviceEmulsion &dv = *em.dev;
    if (first)
        HIP_TRY(launch_first_hit(dv.nodes.as<Node>(), dv.n_nodes, dv.recs.as<float4>(), E->rays_o.as<float>(),
                                 E->rays_d.as<float>(), n_rays, E->hits_i.as<int32_t>(), E->hits_t0.as<float>(),
                                 E->hits_t1.as<float>(), E->stream), FG_ERR_PIPELINE);
    else
        HIP_TRY(launch_all_hits(dv.nodes.as<Node>(), dv.n_nodes, dv.recs.as<float4>(), E->rays_o.as<float>(),
                                E->rays_d.as<float>(), n_rays, max_hits, E->hits_n.as<int32_t>(),
                                E->hits_i.as<int32_t>(), E->hits_t0.as<float>(), E->hits_t1.as<float>(),
                                E->stream), FG_ERR_PIPELINE);
    const size_t m = (size_t)(n_rays * per);
    if (!first)
        HIP_TRY(hipMemcpyAsync(n_hits, E->hits_n.p, (size_t)n_rays * 4, hipMemcpyDeviceToHost, E->stream),
                FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpyAsync(idx, E->hits_i.p, m * 4, hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpyAsync(t0, E->hits_t0.p, m * 4, hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpyAsync(t1, E->hits_t1.p, m * 4, hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    // leaf slots -> crystal indices
    for (int64_t r = 0; r < n_rays; r++) {
        const int64_t nh = first ? 1 : std::min<int64_t>(n_hits[r], max_hits);
        for (int64_t k = 0; k < nh; k++) {
            int32_t &v = idx[r * per + k];
            if (v >= 0) v = (int32_t)em.bvh.idx[(size_t)v];
        }
    }
    return FG_OK;
}

extern "C" fg_status fg_traverse_all_hits(fg_engine *E, const fg_emulsion *em, int64_t n_rays,
                                          const float *o, const float *d, int32_t max_hits,
                                          int32_t *n_hits, int32_t *idx, float *t0, float *t1)
{
    return traverse(E, em, n_rays, o, d, max_hits, n_hits, idx, t0, t1, false);
}

extern "C" fg_status fg_traverse_first_hit(fg_engine *E, const fg_emulsion *em, int64_t n_rays,
                                           const float *o, const float *d, int32_t *idx, float *t0,
                                           float *t1)
{
    return traverse(E, em, n_rays, o, d, 1, nullptr, idx, t0, t1, true);
}

// ------------------------------------------------------------- tile render -------
extern "C" fg_status fg_tile_params_init(fg_tile_params *p, const char *preset, double grain_size_px,
                                         uint32_t tw, uint32_t th, double exposure, double push,
                                         uint64_t tile_seed)
{
    if (!p || !preset || !(grain_size_px > 0))
        return fail(FG_ERR_INVALID_ARG, "fg_tile_params_init: bad argument");
    std::string s(preset);
    for (auto &c : s) c = (char)tolower((unsigned char)c);
    const bool tmax = s == "t-max" || s == "tmax";
    const bool hp5 = s == "hp5" || s == "hp5-plus";
    const bool named1600 = s == "tri-x-1600" || s == "trix-1600" || s == "caffenol";
    // effective push (image_pipeline.rs:368-374)
    const double ep = (s.find("1600") != std::string::npos || s == "caffenol") ? std::max(push, 2.0)
                                                                              : std::max(push, 0.0);
    memset(p, 0, sizeof *p);
    const double px_um = 1.0 / grain_size_px;
    p->film_w = tw * px_um;
    p->film_h = th * px_um;
    p->width_px = tw;
    p->height_px = th;
    p->samples_per_pixel = 4;
    p->seed = tile_seed;
    p->exposure = exposure;
    p->threshold_photons = 4;
    p->max_bounces = 48;
    // emulsion preset as render_tile_cpu_full selects it (image_pipeline.rs:925-938)
    fg_optics o{1.534, 0.003, 0.050, 0.80, 0.14, 0.80};
    if (tmax) {
        p->depth = 8.0; p->r_min = 0.9; p->mean_crystal_radius = 0.40; p->size_variance = 0.20;
        p->morphology = 1; p->tabular_aspect_ratio = 6.5; p->base_sensitivity = 0.8;
        o = {1.534, 0.002, 0.035, 0.85, 0.08, 1.10};
    } else if (hp5) {
        p->depth = 12.5; p->r_min = 1.1; p->mean_crystal_radius = 0.50; p->size_variance = 0.30;
        p->morphology = 2; p->tabular_aspect_ratio = 1.0; p->base_sensitivity = 1.1;
        o = {1.534, 0.0025, 0.045, 0.82, 0.12, 0.85};
    } else if (named1600 || ep >= 1.5) {
        p->depth = 14.5; p->r_min = 1.1; p->mean_crystal_radius = 0.52; p->size_variance = 0.38;
        p->morphology = 0; p->tabular_aspect_ratio = 1.0; p->base_sensitivity = 1.35;
        o = {1.534, 0.0032, 0.052, 0.80, 0.14, 0.80};
    } else {
        p->depth = 14.0; p->r_min = 1.2; p->mean_crystal_radius = 0.55; p->size_variance = 0.40;
        p->morphology = 0; p->tabular_aspect_ratio = 1.0; p->base_sensitivity = 1.2;
    }
    p->film_d = p->depth;
    p->optics = o;
    // development (image_pipeline.rs:963-976)
    p->filament_expansion = tmax ? 1.10 + 0.04 * ep : hp5 ? 1.15 + 0.05 * ep : 1.20 + 0.06 * ep;
    p->clumping_proximity_factor = tmax ? 1.10 + 0.03 * ep : hp5 ? 1.18 + 0.04 * ep : 1.25 + 0.05 * ep;
    // exposure (image_pipeline.rs:942-947) and scan (image_pipeline.rs:996-1002)
    p->photons_per_sq_unit = 6.0 * std::max(exposure, 0.05);
    p->illumination_angle_rad = 0.12;
    p->crystal_absorption_coeff = 1.3;
    p->silver_opacity = 0.25;
    return FG_OK;
}

fg_emulsion_config tile_emulsion_config(const fg_tile_params &p)
{
    fg_emulsion_config c{};
    c.width = p.film_w;
    c.height = p.film_h;
    c.depth = p.depth > 0 ? p.depth : p.film_d;
    c.r_min = p.r_min;
    c.mean_crystal_radius = p.mean_crystal_radius;
    c.size_variance = p.size_variance;
    c.tabular_aspect_ratio = p.tabular_aspect_ratio;
    c.base_sensitivity = p.base_sensitivity;
    c.sensitivity_sigma = 0.15;
    c.min_radius = 0.15;
    c.min_half_thickness = 0.05;
    c.morphology = p.morphology;
    c.k_samples = 30;
    c.optics = p.optics;
    return c;
}

using Clock = std::chrono::steady_clock;
static float ms_since(Clock::time_point t0)
{
    return std::chrono::duration<float, std::milli>(Clock::now() - t0).count();
}

// ------------------------------------------------------ device tile pipeline -----
// Appends one tile to a batched input: its crystal slots from crystal_base on (the leaf
// order of its transport BVH), its transport constants and its TileDev record.  The
// input's pos64 / dims64 / orig arrays must already have room for the slots.
static void append_tile(BatchedInput &in, const Emulsion &em, const fg_exposure_config &x, const fg_tuning &tn,
                        uint32_t x0, uint32_t y0, uint32_t tw, uint32_t th, uint64_t tile_seed,
                        uint32_t node_base, uint32_t n_nodes, int64_t crystal_base, uint64_t out_base,
                        uint32_t out_stride)
{
    if (!em.on_device)
        for (int64_t j = 0; j < em.n; j++) {
            const size_t sl = (size_t)(crystal_base + j);
            const uint32_t ci = em.bvh.idx[(size_t)j];
            for (int a = 0; a < 3; a++) {
                in.pos64[3 * sl + a] = em.pos[3 * (size_t)ci + a];
                in.dims64[3 * sl + a] = em.dims[3 * (size_t)ci + a];
            }
            in.orig[sl] = ci;
        }
    TransportArgs a{};
    fill_transport(a, em.cfg, x, tw, th, tile_seed + 1, em.reach, tn);
    // one launch prunes every tile with one reach: the largest of the tiles' (a larger
    // reach only prunes less, so every tile stays exact); negative = pruning off
    if (in.tiles.empty()) in.transport = a;
    else if (in.transport.reach >= 0.0f) in.transport.reach = std::max(in.transport.reach, a.reach);
    TileDev td{};
    td.x0 = x0;
    td.y0 = y0;
    td.tw = tw;
    td.th = th;
    td.W = a.W;
    td.H = a.H;
    td.pw = a.pw;
    td.ph = a.ph;
    td.key0 = a.key0;
    td.key1 = a.key1;
    td.skey0 = (uint32_t)(tile_seed + 2);
    td.skey1 = (uint32_t)((tile_seed + 2) >> 32);
    td.out_base = out_base;
    td.out_stride = out_stride;
    td.node_base = node_base;
    td.n_nodes = n_nodes;
    td.crystal_base = (uint32_t)crystal_base;
    td.n = (uint32_t)em.n;
    td.n_photons = a.n_photons;
    td.maxb = em.rmax;   // used as given for crystals on the device only
    in.tiles.push_back(td);
}

// Development, scan and transport-occupancy settings shared by a batched input's tiles.
static void batched_settings(BatchedInput &in, const fg_development_config &dc, const fg_scan_config &sc,
                             const fg_tuning &tn, double depth)
{
    in.develop = DevelopParams{dc.threshold_photons, 0u, dc.filament_expansion, dc.clumping_proximity_factor};
    const uint32_t spp = std::max(sc.samples_per_pixel, 1u);
    in.scan.g = std::max((uint32_t)llround(sqrt((double)spp)), 1u);
    in.scan.ss = std::max(sc.supersample, 1u);
    in.scan.mode = sc.mode;
    in.scan.rays_per_lane = tn.scan_rays_per_thread;
    in.scan.z0 = (float)(depth + 0.1);
    in.scan.sigma = (float)sc.silver_opacity;
    in.transport_waves = tn.transport_waves;
}

// Index widths of the batched kernels: 32-bit node ids (two nodes per crystal slot) and
// 32-bit byte offsets into the node and record arrays.
static bool batched_fits(uint64_t n_crystals, uint64_t n_nodes)
{
    return n_crystals * 2 < 0xFFFFFFFFull && offsets_fit_32(n_nodes, n_crystals);
}

// One tile's device arrays, and where its exposure and D_neg live.
struct TileArrays {
    const Node *nodes;
    const Crystal *recs;
    uint32_t n_nodes;
    uint32_t *counts;
    const float *emap;         // exposure map (not NULL when the transport runs), row stride
    uint32_t estride;
    float *dneg;               // D_neg from out_base, row stride out_stride
    uint64_t out_base;
    uint32_t out_stride;
    unsigned long long *stats; // 8 transport counters
    // crystals on the device (frames built there): attributes in crystal order and each
    // leaf slot's crystal index; NULL = the emulsion's host arrays
    const double *dev_pos = nullptr, *dev_dims = nullptr;
    const uint32_t *dev_leaf = nullptr;
};

// A one-tile batched frame (batched.hip) over one tile's device arrays: the per-tile
// paths develop, build the developed-crystal BVH and scan with the batched frame's
// kernels, launched directly (no graph).  *out is NULL on failure.
static fg_status one_tile_batch(fg_engine *E, const Emulsion &em, const TileArrays &d, uint32_t x0, uint32_t y0,
                                uint32_t tw, uint32_t th, uint64_t tile_seed, const fg_exposure_config &x,
                                const fg_development_config &dc, const fg_scan_config &sc, BatchedFrame **out)
{
    *out = nullptr;
    if (!batched_fits((uint64_t)em.n, d.n_nodes))
        return fail(FG_ERR_INVALID_ARG, "tile: too many crystals for the device pipeline");
    BatchedInput in;
    in.nodes = d.nodes;
    in.recs = d.recs;
    in.counts = d.counts;
    in.emap = d.emap;
    in.dneg = d.dneg;
    in.stats = d.stats;
    in.frame_w = d.estride;
    in.frame_h = th;
    in.n_crystals = em.n;
    in.dev_pos = d.dev_pos;
    in.dev_dims = d.dev_dims;
    in.dev_leaf = d.dev_leaf;
    if (!in.dev_pos) {
        in.pos64.resize((size_t)em.n * 3);
        in.dims64.resize((size_t)em.n * 3);
        in.orig.resize((size_t)em.n);
    }
    append_tile(in, em, x, E->tuning, x0, y0, tw, th, tile_seed, 0u, d.n_nodes, 0, d.out_base, d.out_stride);
    batched_settings(in, dc, sc, E->tuning, em.cfg.depth);
    BatchedFrame *b = nullptr;
    const hipError_t e = batched_create(in, E->stream, &b);
    if (e != hipSuccess) {
        batched_destroy(b);
        return fail(FG_ERR_BUFFER_RANGE, std::string("tile pipeline: ") + hipGetErrorString(e));
    }
    *out = b;
    return FG_OK;
}

struct BatchGuard {
    BatchedFrame *b = nullptr;
    ~BatchGuard() { batched_destroy(b); }
};

// develop_emulsion (development.rs:59-188) of one emulsion on the device: the batched
// frame's threshold, union-find clumping and boost kernels over a one-tile frame, from
// the emulsion's latent counts; flags, expansions and clump ids come back to the host.
static fg_status develop_device(fg_engine *E, Emulsion &em, const fg_development_config &cfg,
                                fg_development_report &rep)
{
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    fg_status s = ensure_device(E, em);
    if (s != FG_OK) return s;
    const size_t n = (size_t)em.n;
    std::vector<uint32_t> slots(n);
    for (size_t k = 0; k < n; k++) slots[k] = em.counts[em.bvh.idx[k]];
    const DeviceEmulsion &d = *em.dev;
    HIP_TRY(hipMemcpyAsync(d.counts.p, slots.data(), n * 4, hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    const TileArrays ta{d.nodes.as<Node>(), d.recs.as<Crystal>(), d.n_nodes, d.counts.as<uint32_t>(), nullptr, 1,
                        nullptr, 0, 1, nullptr};
    const fg_exposure_config x{};
    const fg_scan_config sc{1, 1, 1, FG_SCAN_DNEG, 0.0, 1, 0};
    BatchGuard bg;
    s = one_tile_batch(E, em, ta, 0, 0, 1, 1, 0, x, cfg, sc, &bg.b);
    if (s != FG_OK) return s;
    HIP_TRY(batched_develop(bg.b, E->stream), FG_ERR_PIPELINE);
    std::vector<TileDev> tiles;
    uint32_t err[4];
    HIP_TRY(batched_read_tiles(bg.b, E->stream, tiles, err), FG_ERR_BUFFER_MAP);
    std::vector<uint8_t> dv;
    std::vector<double> fe;
    std::vector<int32_t> cl;
    HIP_TRY(batched_read_develop(bg.b, E->stream, dv, fe, cl), FG_ERR_BUFFER_MAP);
    for (size_t k = 0; k < n; k++) {
        const uint32_t i = em.bvh.idx[k];
        em.developed[i] = dv[k];
        em.fe[i] = fe[k];
        em.clump[i] = cl[k];
    }
    const TileDev &t = tiles[0];
    memset(&rep, 0, sizeof rep);
    rep.total_crystals = n;
    rep.developed_crystals = t.n_dev;
    rep.fixed_crystals = n - t.n_dev;
    rep.total_clumps = t.clumps;
    rep.max_clump_size = t.max_clump;
    rep.avg_clump_size = t.clumps > 0 ? (double)t.n_dev / (double)t.clumps : 0.0;
    return FG_OK;
}

extern "C" fg_status fg_render_tile_all_gpu(fg_engine *E, const float *tile_img, const fg_tile_params *p,
                                            float *out, fg_tile_stats *stats)
{
    if (!E || !p) return fail(FG_ERR_INVALID_ARG, "fg_render_tile_all_gpu: null argument");
    if ((uint64_t)p->width_px * p->height_px == 0) return FG_OK;
    if (!tile_img || !out) return fail(FG_ERR_INVALID_ARG, "fg_render_tile_all_gpu: null buffer");
    const fg_emulsion_config ec = tile_emulsion_config(*p);
    std::string msg;
    if (!valid_cfg(&ec, msg)) return fail(FG_ERR_INVALID_ARG, "fg_render_tile_all_gpu: " + msg);
    // Synthesis runs on the calling thread outside the engine lock (the reference's
    // dispatch_lock, gpu.rs:169, 503, serialises GPU work only), so callers on several
    // threads overlap it; transport, development, the developed-crystal BVH and the
    // scan run on the device as a one-tile batched frame.
    const auto t0 = Clock::now();
    fg_emulsion em;
    synthesize(ec, p->seed, em.em);
    if (!finalize_emulsion(em.em, msg)) return fail(FG_ERR_INVALID_ARG, "fg_render_tile_all_gpu: " + msg);
    fg_exposure_config xc{};
    xc.photons_per_sq_unit = p->photons_per_sq_unit;
    xc.illumination_angle_rad = p->illumination_angle_rad;
    xc.crystal_absorption_coeff = p->crystal_absorption_coeff;
    xc.max_bounces = p->max_bounces;
    xc.photons_per_pixel = 0;
    const fg_development_config dc{p->threshold_photons, 0, p->filament_expansion, p->clumping_proximity_factor, 0.85};
    const fg_scan_config sc{p->width_px, p->height_px, p->samples_per_pixel, FG_SCAN_DNEG, p->silver_opacity, 1, 0};
    const uint32_t tw = p->width_px, th = p->height_px;
    const size_t npx = (size_t)tw * th;
    std::lock_guard<std::mutex> lk(E->mu);
    struct DeviceGuard {
        Emulsion &em;
        ~DeviceGuard() { free_device(em); }
    } dguard{em.em};
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    fg_status s = ensure_device(E, em.em);
    if (s != FG_OK) return s;
    HIP_TRY(E->emap.ensure(npx * sizeof(float)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->out.ensure(npx * sizeof(float)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->stats.ensure(8 * sizeof(unsigned long long)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemcpyAsync(E->emap.p, tile_img, npx * sizeof(float), hipMemcpyHostToDevice, E->stream),
            FG_ERR_BUFFER_MAP);
    const DeviceEmulsion &d = *em.em.dev;
    const TileArrays ta{d.nodes.as<Node>(), d.recs.as<Crystal>(), d.n_nodes, d.counts.as<uint32_t>(),
                        E->emap.as<float>(), tw, E->out.as<float>(), 0, tw, E->stats.as<unsigned long long>()};
    BatchGuard bg;
    s = one_tile_batch(E, em.em, ta, 0, 0, tw, th, p->seed, xc, dc, sc, &bg.b);
    if (s != FG_OK) return s;
    HIP_TRY(batched_render(bg.b, E->stream, false), FG_ERR_PIPELINE);
    HIP_TRY(hipMemcpyAsync(out, E->out.p, npx * sizeof(float), hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    std::vector<TileDev> tiles;
    uint32_t err[4];
    HIP_TRY(batched_read_tiles(bg.b, E->stream, tiles, err), FG_ERR_BUFFER_MAP);
    if (err[0] || err[1] || err[2])
        return fail(FG_ERR_PIPELINE, "fg_render_tile_all_gpu: developed-crystal BVH exceeds the reference stack depth or leaf size");
    if (stats) {
        float tk = 0, dk = 0, sk = 0;
        HIP_TRY(batched_kernel_ms(bg.b, &tk, &dk, &sk), FG_ERR_CHANNEL);
        stats->crystals = (uint64_t)em.em.n;
        stats->developed = tiles[0].n_dev;
        stats->clumps = tiles[0].clumps;
        stats->photons = tiles[0].n_photons;
        stats->ms_transport = tk;
        stats->ms_develop = dk;
        stats->ms_scan = sk;
        stats->ms_total = ms_since(t0);
    }
    return FG_OK;
}

extern "C" fg_status fg_render_tile_delta_d(fg_engine *E, int64_t n, const double *pos, const double *dims,
                                            const double *quat, const int32_t *morph, const double *fe,
                                            const fg_tile_params *p, float *out_delta_d, float *out_dneg)
{
    if (!E || !p || n < 0 || (n > 0 && (!pos || !dims || !quat || !morph || !fe)))
        return fail(FG_ERR_INVALID_ARG, "fg_render_tile_delta_d: null or negative argument");
    if ((uint64_t)p->width_px * p->height_px == 0) return FG_OK;
    if (!out_delta_d) return fail(FG_ERR_INVALID_ARG, "fg_render_tile_delta_d: null output");
    if (n > (int64_t)(1u << 27)) return fail(FG_ERR_INVALID_ARG, "fg_render_tile_delta_d: too many crystals");
    if (!(p->film_w > 0) || !(p->film_h > 0) || !(p->film_d > 0) || !std::isfinite(p->film_w) ||
        !std::isfinite(p->film_h) || !std::isfinite(p->film_d) || !std::isfinite(p->silver_opacity))
        return fail(FG_ERR_INVALID_ARG, "fg_render_tile_delta_d: film extents must be positive and finite");
    const fg_scan_config sc{p->width_px, p->height_px, p->samples_per_pixel, FG_SCAN_DNEG, p->silver_opacity, 1, 0};
    std::string msg;
    if (!valid_scan(&sc, msg)) return fail(FG_ERR_INVALID_ARG, "fg_render_tile_delta_d: " + msg);
    // every given crystal is developed with its own filament expansion (the caller passes
    // the developed ones, gpu.rs:941-944); the scan builds its BVH over them
    Emulsion em;
    em.cfg = tile_emulsion_config(*p);
    em.cfg.depth = p->film_d;
    em.n = n;
    em.pos.assign(pos, pos + 3 * n);
    em.dims.assign(dims, dims + 3 * n);
    em.quat.assign(quat, quat + 4 * n);
    em.morph.assign(morph, morph + n);
    em.sens.assign((size_t)n, 1.0);
    em.developed.assign((size_t)n, 1u);
    em.fe.assign(fe, fe + n);
    for (int64_t i = 0; i < n; i++) {
        if (morph[i] < 0 || morph[i] > 2)
            return fail(FG_ERR_INVALID_ARG, "fg_render_tile_delta_d: bad morphology");
        if (!(fe[i] > 0) || !std::isfinite(fe[i]))
            return fail(FG_ERR_INVALID_ARG, "fg_render_tile_delta_d: filament expansion must be positive");
    }
    ScanBvh sb;
    if (!build_scan_bvh(em, sb, msg)) return fail(FG_ERR_INVALID_ARG, "fg_render_tile_delta_d: " + msg);
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    const float *dneg = nullptr;
    fg_status s = scan_enqueue(E, em, sb, sc, p->seed, &dneg);
    if (s != FG_OK) return s;
    const size_t npx = (size_t)p->width_px * p->height_px;
    HIP_TRY(E->hp.ensure(npx * sizeof(float)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(launch_highpass17(dneg, E->hp.as<float>(), p->width_px, p->height_px, E->stream), FG_ERR_PIPELINE);
    HIP_TRY(hipMemcpyAsync(out_delta_d, E->hp.p, npx * sizeof(float), hipMemcpyDeviceToHost, E->stream),
            FG_ERR_BUFFER_MAP);
    if (out_dneg)
        HIP_TRY(hipMemcpyAsync(out_dneg, dneg, npx * sizeof(float), hipMemcpyDeviceToHost, E->stream),
                FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    return FG_OK;
}

// -------------------------------------------------------------------- frames -----
struct TileRec {
    uint32_t x0, y0, tw, th;   // first pixel and clipped extent
    uint64_t seed;
    int64_t crystal_base, node_base;
    uint32_t n_nodes;
    int64_t n;
    int64_t out_base;          // first D_neg value of the tile
    uint32_t out_stride;       // D_neg row stride
};

struct fg_frame {
    fg_frame_desc desc{};
    std::vector<float> emap;
    std::vector<TileRec> tiles;
    int64_t dneg_values = 0;   // width * height, or the sum of tile areas with overlap
    fg::FilmLayout layout{};   // tile spans and Hann ramps of an overlapped frame
    DevBuf d_layout;
    std::vector<Emulsion> ems;
    int64_t n_crystals = 0, n_nodes = 0;
    DevBuf d_nodes, d_recs, d_counts, d_emap, d_dneg, d_dneg2, d_stats;
    bool rendered = false;
    bool otf_applied = false;
    uint32_t last_mode = FG_FRAME_PLAIN;
    BatchedFrame *batched = nullptr;
    bool graph_failed = false;
    fg_tuning batched_tuning{};
    std::vector<BatchedFrame *> plain;   // one-tile frames of the plain path, per tile
    fg_tuning plain_tuning{};
    // frames built on the device: crystal attributes (crystal order, tile t's from its
    // crystal_base) and each leaf slot's crystal index within its tile
    bool device_built = false;
    DevBuf d_pos, d_dims, d_quat, d_sens, d_leaf;
};

extern "C" int64_t fg_frame_crystal_count(const fg_frame *F) { return F ? F->n_crystals : 0; }

// Emulsions of a device-synthesised frame built on the device: dart-thrown centres of
// every tile at once (synth.hip), compacted in cell order, each crystal's attributes
// from its own Philox stream (emulsion.rs:160-207), and every tile's transport BVH
// (BVH::build, bvh.rs:157-295) and leaf-order records written straight into the frame's
// node and record arrays (tile k's nodes from 2 crystal_base).  The host keeps each
// tile's crystals and leaf order (F->ems) for the batched input and the read-backs; no
// BVH is built on the host.  Caller holds E->mu.
static fg_status device_frame_emulsions(fg_engine *E, fg_frame *F, const std::vector<fg_emulsion_config> &cfgs,
                                        const std::vector<uint64_t> &seeds)
{
    const size_t nt = F->tiles.size();
    std::vector<size_t> off;
    fg_status st = device_cells_tiles(E, cfgs, seeds, nullptr, off);
    if (st != FG_OK) return st;
    std::vector<SynthTile> tiles(nt);
    for (size_t k = 0; k < nt; k++) {
        tiles[k] = SynthTile{};
        tiles[k].cell_off = off[k];
        tiles[k].n_cells = (uint32_t)(off[k + 1] - off[k]);
        tiles[k].seed = seeds[k];
        tiles[k].rmax = cfgs[k].r_min * 0.48;
    }
    DevBuf dtiles, dcounts;
    HIP_TRY(dtiles.ensure(nt * sizeof(SynthTile)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(dcounts.ensure(nt * sizeof(uint32_t)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemcpyAsync(dtiles.p, tiles.data(), nt * sizeof(SynthTile), hipMemcpyHostToDevice, E->stream),
            FG_ERR_BUFFER_MAP);
    HIP_TRY(launch_synth_count(E->synth_cells.as<float4>(), dtiles.as<SynthTile>(), (uint32_t)nt,
                               dcounts.as<uint32_t>(), E->stream), FG_ERR_PIPELINE);
    std::vector<uint32_t> counts(nt);
    HIP_TRY(hipMemcpyAsync(counts.data(), dcounts.p, nt * sizeof(uint32_t), hipMemcpyDeviceToHost, E->stream),
            FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    uint64_t n = 0;
    TransportBuildIn in;
    for (size_t k = 0; k < nt; k++) {
        if (n + counts[k] >= (1ull << 31)) return fail(FG_ERR_INVALID_ARG, "device synthesis: too many crystals");
        tiles[k].crystal_base = (uint32_t)n;
        in.base.push_back((uint32_t)n);
        in.count.push_back(counts[k]);
        n += counts[k];
    }
    if (!batched_fits(n, 2 * n)) return fail(FG_ERR_INVALID_ARG, "device synthesis: node or record arrays of 4 GiB or more");
    HIP_TRY(hipMemcpyAsync(dtiles.p, tiles.data(), nt * sizeof(SynthTile), hipMemcpyHostToDevice, E->stream),
            FG_ERR_BUFFER_MAP);
    DevBuf &pos = F->d_pos, &dims = F->d_dims, &quat = F->d_quat, &sens = F->d_sens, &leaf = F->d_leaf;
    HIP_TRY(pos.ensure(3 * n * sizeof(double)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(dims.ensure(3 * n * sizeof(double)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(quat.ensure(4 * n * sizeof(double)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(sens.ensure(n * sizeof(double)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(leaf.ensure(n * sizeof(uint32_t)), FG_ERR_BUFFER_RANGE);
    const fg_emulsion_config &c0 = cfgs.empty() ? F->desc.emulsion : cfgs[0];
    SynthAttrs A{};
    A.mu = log(c0.mean_crystal_radius);
    A.size_sd = c0.size_variance;
    A.min_radius = c0.min_radius;
    A.base_sens = c0.base_sensitivity;
    A.sens_sd = c0.sensitivity_sigma;
    A.aspect = c0.tabular_aspect_ratio;
    A.min_half_thickness = c0.min_half_thickness;
    A.morph = c0.morphology;
    HIP_TRY(launch_synth_fill(E->synth_cells.as<float4>(), dtiles.as<SynthTile>(), (uint32_t)nt, A, pos.as<double>(),
                              dims.as<double>(), quat.as<double>(), sens.as<double>(), E->stream), FG_ERR_PIPELINE);
    HIP_TRY(F->d_nodes.ensure(2 * n * sizeof(Node)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(F->d_recs.ensure(n * sizeof(Crystal)), FG_ERR_BUFFER_RANGE);
    in.pos = pos.as<double>();
    in.dims = dims.as<double>();
    in.quat = quat.as<double>();
    in.sens = sens.as<double>();
    in.morph = c0.morphology;
    in.n = (int64_t)n;
    std::vector<uint32_t> node_count;
    uint32_t err[4];
    HIP_TRY(build_transport_bvhs(in, F->d_nodes.as<Node>(), F->d_recs.as<Crystal>(), leaf.as<uint32_t>(), node_count,
                                 err, E->stream), FG_ERR_PIPELINE);
    if (err[0] || err[1] || err[2])
        return fail(FG_ERR_PIPELINE, "device synthesis: transport BVH exceeds the reference stack depth or leaf size");
    // the host keeps each tile's leaf order (for the read-backs); the attributes stay on
    // the device, where the batched inputs gather them
    for (size_t k = 0; k < nt; k++) {
        TileRec &t = F->tiles[k];
        Emulsion &em = F->ems[k];
        const size_t b = in.base[k], m = in.count[k];
        em.cfg = cfgs[k];
        em.n = (int64_t)m;
        em.on_device = true;
        em.morph.assign(m, (int32_t)c0.morphology);
        em.counts.assign(m, 0u);
        em.developed.assign(m, 0u);
        em.fe.assign(m, 1.0);
        em.clump.assign(m, -1);
        em.bvh = RefBvh{};
        em.bvh.idx.resize(m);
        if (m)
            HIP_TRY(hipMemcpyAsync(em.bvh.idx.data(), leaf.as<uint32_t>() + b, m * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
        // the radius clamp r_min * 0.48 (emulsion.rs:166) bounds every semi-axis: with it
        // the transport prunes by a reach no smaller than the crystals' own, so exactly
        em.rmax = tiles[k].rmax;
        em.reach = c0.morphology == 0 ? (sqrt(3.0) - 1.0) * em.rmax : 0.0;
        t.crystal_base = (int64_t)b;
        t.n = (int64_t)m;
        t.node_base = 2 * (int64_t)b;
        t.n_nodes = node_count[k];
    }
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    F->device_built = true;
    F->n_crystals = (int64_t)n;
    F->n_nodes = 2 * (int64_t)n;
    return FG_OK;
}

extern "C" fg_status fg_frame_create(fg_engine *E, const fg_frame_desc *desc, fg_frame **out)
{
    if (!E || !desc || !out || !desc->exposure_map)
        return fail(FG_ERR_INVALID_ARG, "fg_frame_create: null argument");
    if (desc->width == 0 || desc->height == 0 || desc->tile == 0 || !(desc->pixel_size > 0))
        return fail(FG_ERR_INVALID_ARG, "fg_frame_create: empty frame or tile");
    if (desc->overlap && (2ull * desc->overlap > desc->tile || desc->scan.otf))
        return fail(FG_ERR_INVALID_ARG, "fg_frame_create: overlap needs 2 * overlap <= tile and scan.otf = 0");
    if ((uint64_t)desc->width * desc->height > (1ull << 31))
        return fail(FG_ERR_INVALID_ARG, "fg_frame_create: frame above 2^31 pixels");
    std::string msg;
    if (!valid_exposure(&desc->exposure, msg) || !valid_scan(&desc->scan, msg))
        return fail(FG_ERR_INVALID_ARG, "fg_frame_create: " + msg);
    fg_frame *F = new fg_frame();
    F->desc = *desc;
    F->emap.assign(desc->exposure_map, desc->exposure_map + (size_t)desc->width * desc->height);
    F->desc.exposure_map = nullptr;
    F->desc.density_mult = nullptr;
    const uint32_t nx = (desc->width + desc->tile - 1) / desc->tile;
    const uint32_t ny = (desc->height + desc->tile - 1) / desc->tile;
    std::vector<double> mult((size_t)nx * ny, 1.0);
    if (desc->density_mult)
        for (size_t k = 0; k < mult.size(); k++) mult[k] = desc->density_mult[k];
    const uint32_t ov = desc->overlap;
    for (uint32_t ty = 0; ty < ny; ty++)
        for (uint32_t tx = 0; tx < nx; tx++) {
            TileRec t{};
            // tile span (image_pipeline.rs:400-410); abutting tiles when the overlap is 0
            fg::film_span(tx, desc->tile, ov, desc->width, t.x0, t.tw);
            fg::film_span(ty, desc->tile, ov, desc->height, t.y0, t.th);
            const uint64_t idx = (uint64_t)ty * nx + tx;
            // tile seed (image_pipeline.rs:430-434)
            t.seed = desc->seed + idx * 73856093ull + (uint64_t)tx * 19349663ull + (uint64_t)ty * 83492791ull;
            if (ov) {
                t.out_base = F->dneg_values;
                t.out_stride = t.tw;
                F->dneg_values += (int64_t)t.tw * t.th;
            } else {
                t.out_base = (int64_t)t.y0 * desc->width + t.x0;
                t.out_stride = desc->width;
            }
            F->tiles.push_back(t);
        }
    if (!ov) F->dneg_values = (int64_t)desc->width * desc->height;
    F->ems.resize(F->tiles.size());
    auto tile_cfg = [&](size_t k) {
        const TileRec &t = F->tiles[k];
        fg_emulsion_config c = desc->emulsion;
        c.width = t.tw * desc->pixel_size;
        c.height = t.th * desc->pixel_size;
        c.r_min = desc->emulsion.r_min / cbrt(std::max(mult[k], 1e-3));
        return c;
    };
    if (desc->synthesis > 1) {
        delete F;
        return fail(FG_ERR_INVALID_ARG, "fg_frame_create: synthesis must be 0 (host) or 1 (device)");
    }
    const bool device_built = desc->synthesis == 1;
    if (device_built) {
        // every tile's emulsion and transport BVH on the device (device_frame_emulsions)
        std::vector<fg_emulsion_config> cfgs;
        std::vector<uint64_t> seeds;
        for (size_t k = 0; k < F->tiles.size(); k++) {
            cfgs.push_back(tile_cfg(k));
            seeds.push_back(F->tiles[k].seed);
        }
        std::lock_guard<std::mutex> lk(E->mu);
        hipError_t e = hipSetDevice(E->device);
        if (e != hipSuccess) {
            delete F;
            return fail(FG_ERR_DEVICE_REQUEST, std::string("hipSetDevice: ") + hipGetErrorString(e));
        }
        fg_status s = device_frame_emulsions(E, F, cfgs, seeds);
        if (s != FG_OK) {
            fg_frame_destroy(F);
            return s;
        }
    }
    if (!device_built) {
        // the reference's Bridson sampler and BVH per tile on the host threads
        std::atomic<size_t> next{0};
        std::atomic<bool> bad{false};
        std::string err;
        std::mutex emu;
        unsigned nt = desc->threads ? desc->threads : std::max(1u, std::thread::hardware_concurrency());
        nt = std::min<unsigned>(nt, (unsigned)F->tiles.size());
        auto work = [&]() {
            for (;;) {
                const size_t k = next.fetch_add(1);
                if (k >= F->tiles.size()) break;
                synthesize(tile_cfg(k), F->tiles[k].seed, F->ems[k]);
                std::string m;
                if (!finalize_emulsion(F->ems[k], m)) {
                    std::lock_guard<std::mutex> g(emu);
                    bad = true;
                    err = m;
                }
            }
        };
        std::vector<std::thread> pool;
        for (unsigned i = 0; i < nt; i++) pool.emplace_back(work);
        for (auto &th : pool) th.join();
        if (bad) {
            delete F;
            return fail(FG_ERR_INVALID_ARG, "fg_frame_create: " + err);
        }
        for (size_t k = 0; k < F->tiles.size(); k++) {
            TileRec &t = F->tiles[k];
            t.crystal_base = F->n_crystals;
            t.node_base = F->n_nodes;
            t.n = F->ems[k].n;
            t.n_nodes = (uint32_t)F->ems[k].nodes.size();
            F->n_crystals += t.n;
            F->n_nodes += t.n_nodes;
        }
    }
    std::lock_guard<std::mutex> lk(E->mu);
    auto cleanup = [&](fg_status s) { fg_frame_destroy(F); return s; };
    hipError_t e = hipSetDevice(E->device);
    if (e == hipSuccess) e = F->d_nodes.ensure((size_t)F->n_nodes * sizeof(Node));
    if (e == hipSuccess) e = F->d_recs.ensure((size_t)F->n_crystals * sizeof(Crystal));
    if (e == hipSuccess) e = F->d_counts.ensure((size_t)F->n_crystals * 4);
    if (e == hipSuccess) e = F->d_emap.ensure(F->emap.size() * 4);
    if (e == hipSuccess) e = F->d_dneg.ensure((size_t)F->dneg_values * 4);
    if (e == hipSuccess) e = F->d_dneg2.ensure(F->emap.size() * 4);
    if (e == hipSuccess && ov)
        e = fg::film_layout_upload(desc->width, desc->height, desc->tile, ov, F->d_layout, F->layout, E->stream);
    if (e == hipSuccess) e = F->d_stats.ensure(64);
    if (e != hipSuccess) return cleanup(fail(FG_ERR_BUFFER_RANGE, std::string("frame allocation: ") + hipGetErrorString(e)));
    for (size_t k = 0; k < F->tiles.size() && e == hipSuccess && !device_built; k++) {
        const TileRec &t = F->tiles[k];
        if (t.n_nodes)
            e = hipMemcpy(F->d_nodes.as<Node>() + t.node_base, F->ems[k].nodes.data(), t.n_nodes * sizeof(Node),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess && t.n)
            e = hipMemcpy(F->d_recs.as<Crystal>() + t.crystal_base, F->ems[k].recs.data(),
                          (size_t)t.n * sizeof(Crystal), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = hipMemcpy(F->d_emap.p, F->emap.data(), F->emap.size() * 4, hipMemcpyHostToDevice);
    // the uploads above run on the null stream, which does not order the engine's
    // non-blocking stream: wait for them before any kernel reads the frame
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return cleanup(fail(FG_ERR_BUFFER_MAP, std::string("frame upload: ") + hipGetErrorString(e)));
    *out = F;
    return FG_OK;
}

extern "C" void fg_frame_destroy(fg_frame *F)
{
    if (!F) return;
    if (F->batched) batched_destroy(F->batched);
    for (BatchedFrame *b : F->plain) batched_destroy(b);
    for (DevBuf *b : {&F->d_nodes, &F->d_recs, &F->d_counts, &F->d_emap, &F->d_dneg, &F->d_dneg2,
                      &F->d_stats, &F->d_layout, &F->d_pos, &F->d_dims, &F->d_quat, &F->d_sens, &F->d_leaf})
        b->release();
    delete F;
}

// Plain per-tile path: per tile one launch sequence (transport; development and the
// developed-crystal BVH; scan -- a one-tile batched frame launched directly, no graph)
// and a synchronisation with the readback of the tile's counters: the shape of
// render_tile_all_gpu called per tile, with the frame's inputs resident.  The one-tile
// frames are built on the first plain render and kept with the frame.
static fg_status frame_render_plain(fg_engine *E, fg_frame *F, fg_frame_stats *st)
{
    const fg_frame_desc &D = F->desc;
    if (!F->plain.empty() && memcmp(&F->plain_tuning, &E->tuning, sizeof(fg_tuning)) != 0) {
        for (BatchedFrame *b : F->plain) batched_destroy(b);
        F->plain.clear();
    }
    if (F->plain.empty()) {
        F->plain_tuning = E->tuning;
        F->plain.assign(F->tiles.size(), nullptr);
        for (size_t k = 0; k < F->tiles.size(); k++) {
            const TileRec &t = F->tiles[k];
            TileArrays ta{F->d_nodes.as<Node>() + t.node_base, F->d_recs.as<Crystal>() + t.crystal_base,
                          t.n_nodes, F->d_counts.as<uint32_t>() + t.crystal_base, F->d_emap.as<float>(),
                          D.width, F->d_dneg.as<float>(), (uint64_t)t.out_base, t.out_stride,
                          F->d_stats.as<unsigned long long>()};
            if (F->device_built) {
                ta.dev_pos = F->d_pos.as<double>() + 3 * t.crystal_base;
                ta.dev_dims = F->d_dims.as<double>() + 3 * t.crystal_base;
                ta.dev_leaf = F->d_leaf.as<uint32_t>() + t.crystal_base;
            }
            fg_status s = one_tile_batch(E, F->ems[k], ta, t.x0, t.y0, t.tw, t.th, t.seed, D.exposure,
                                         D.development, D.scan, &F->plain[k]);
            if (s != FG_OK) return s;
        }
    }
    const uint32_t spp = std::max(D.scan.samples_per_pixel, 1u);
    const uint64_t g = std::max((uint32_t)llround(sqrt((double)spp)), 1u), ss = std::max(D.scan.supersample, 1u);
    unsigned long long tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    float ms_tk = 0, ms_dk = 0, ms_sk = 0;
    for (size_t k = 0; k < F->tiles.size(); k++) {
        BatchedFrame *b = F->plain[k];
        HIP_TRY(batched_render(b, E->stream, false), FG_ERR_PIPELINE);
        unsigned long long cnt[8];
        HIP_TRY(hipMemcpyAsync(cnt, F->d_stats.p, sizeof cnt, hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
        std::vector<TileDev> tiles;
        uint32_t err[4];
        HIP_TRY(batched_read_tiles(b, E->stream, tiles, err), FG_ERR_BUFFER_MAP);   // synchronises
        if (err[0] || err[1] || err[2])
            return fail(FG_ERR_PIPELINE, "plain frame: developed-crystal BVH exceeds the reference stack depth or leaf size");
        for (int i = 0; i < 8; i++) tot[i] += cnt[i];
        float tk = 0, dk = 0, sk = 0;
        HIP_TRY(batched_kernel_ms(b, &tk, &dk, &sk), FG_ERR_CHANNEL);
        ms_tk += tk;
        ms_dk += dk;
        ms_sk += sk;
        const TileDev &t = tiles[0];
        st->photons += t.n_photons;
        st->developed += t.n_dev;
        st->clumps += t.clumps;
        st->rays += (uint64_t)t.tw * t.th * g * g * ss * ss;
    }
    st->bounces = tot[0];
    st->nodes_visited = tot[1];
    st->prims_tested = tot[2];
    st->abs_events = tot[3];
    st->atomics = tot[5];
    st->packets_run = tot[7];
    st->ms_kernel_transport = ms_tk;
    st->ms_kernel_scan = ms_sk;
    st->ms_transport = ms_tk;
    st->ms_develop = ms_dk;
    st->ms_scan = ms_sk;
    return FG_OK;
}

// Batched path: one device pass over all tiles (batched.hip), replayed as a hipGraph.
static fg_status frame_render_batched(fg_engine *E, fg_frame *F, fg_frame_stats *st)
{
    const fg_frame_desc &D = F->desc;
    if (F->batched && memcmp(&F->batched_tuning, &E->tuning, sizeof(fg_tuning)) != 0) {
        batched_destroy(F->batched);  // launch parameters are baked into the graph
        F->batched = nullptr;
    }
    if (!F->batched) {
        F->batched_tuning = E->tuning;
        BatchedInput in;
        in.nodes = F->d_nodes.as<Node>();
        in.recs = F->d_recs.as<Crystal>();
        in.counts = F->d_counts.as<uint32_t>();
        in.emap = F->d_emap.as<float>();
        in.dneg = F->d_dneg.as<float>();
        in.stats = F->d_stats.as<unsigned long long>();
        in.frame_w = D.width;
        in.frame_h = D.height;
        in.n_crystals = F->n_crystals;
        if (F->device_built) {
            in.dev_pos = F->d_pos.as<double>();
            in.dev_dims = F->d_dims.as<double>();
            in.dev_leaf = F->d_leaf.as<uint32_t>();
        } else {
            in.pos64.resize((size_t)F->n_crystals * 3);
            in.dims64.resize((size_t)F->n_crystals * 3);
            in.orig.resize((size_t)F->n_crystals);
        }
        for (size_t k = 0; k < F->tiles.size(); k++) {
            const TileRec &t = F->tiles[k];
            append_tile(in, F->ems[k], D.exposure, E->tuning, t.x0, t.y0, t.tw, t.th, t.seed, (uint32_t)t.node_base,
                        t.n_nodes, t.crystal_base, (uint64_t)t.out_base, t.out_stride);
        }
        batched_settings(in, D.development, D.scan, E->tuning, D.emulsion.depth);
        if (!batched_fits((uint64_t)F->n_crystals, (uint64_t)F->n_nodes))
            return fail(FG_ERR_INVALID_ARG, "batched frame: too many crystals for 32-bit node ids and offsets");
        HIP_TRY(batched_create(in, E->stream, &F->batched), FG_ERR_BUFFER_RANGE);
    }
    HIP_TRY(hipEventRecord(E->ev0, E->stream), FG_ERR_CHANNEL);
    hipError_t e = batched_render(F->batched, E->stream, !F->graph_failed);
    if (e != hipSuccess && !F->graph_failed) {
        // graph capture unavailable: launch the same sequence directly from now on, and say
        // so once (fg_frame_stats.graph_used reports it per render)
        F->graph_failed = true;
        fprintf(stderr, "fgrain_hip: hipGraph of the batched frame unavailable (%s: %s); "
                        "launching its kernels directly\n",
                batched_graph_failure(F->batched), hipGetErrorString(e));
        (void)hipGetLastError();
        e = batched_render(F->batched, E->stream, false);
    }
    HIP_TRY(e, FG_ERR_PIPELINE);
    st->graph_used = F->graph_failed ? 0u : 1u;
    HIP_TRY(hipEventRecord(E->ev1, E->stream), FG_ERR_CHANNEL);
    std::vector<TileDev> tiles;
    uint32_t err[4];
    HIP_TRY(batched_read_tiles(F->batched, E->stream, tiles, err), FG_ERR_BUFFER_MAP);
    if (err[0] || err[1] || err[2])
        return fail(FG_ERR_PIPELINE, "batched frame: developed-crystal BVH exceeds the reference stack depth or leaf size");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, E->ev0, E->ev1);
    unsigned long long cnt[8];
    HIP_TRY(hipMemcpy(cnt, F->d_stats.p, 64, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    st->bounces = cnt[0];
    st->nodes_visited = cnt[1];
    st->prims_tested = cnt[2];
    st->abs_events = cnt[3];
    st->atomics = cnt[5];
    st->packets_run = cnt[7];
    const uint32_t spp = std::max(D.scan.samples_per_pixel, 1u);
    const uint64_t g = std::max((uint32_t)llround(sqrt((double)spp)), 1u), ss = std::max(D.scan.supersample, 1u);
    for (const TileDev &t : tiles) {
        st->photons += t.n_photons;
        st->developed += t.n_dev;
        st->clumps += t.clumps;
        st->rays += (uint64_t)t.tw * t.th * g * g * ss * ss;
    }
    float tk = 0, dk = 0, sk = 0;
    HIP_TRY(batched_kernel_ms(F->batched, &tk, &dk, &sk), FG_ERR_CHANNEL);
    st->ms_kernel_transport = tk;
    st->ms_kernel_scan = sk;
    st->ms_transport = tk;
    st->ms_develop = dk;
    st->ms_scan = sk;
    (void)ms;
    return FG_OK;
}

extern "C" fg_status fg_frame_render(fg_engine *E, fg_frame *F, uint32_t mode, fg_frame_stats *stats)
{
    if (!E || !F) return fail(FG_ERR_INVALID_ARG, "fg_frame_render: null argument");
    if (mode != FG_FRAME_PLAIN && mode != FG_FRAME_BATCHED)
        return fail(FG_ERR_INVALID_ARG, "fg_frame_render: unknown mode");
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    fg_frame_stats st{};
    st.tiles = F->tiles.size();
    st.crystals = (uint64_t)F->n_crystals;
    st.pixels = (uint64_t)F->desc.width * F->desc.height;
    const auto t0 = Clock::now();
    fg_status s;
    if (mode == FG_FRAME_PLAIN) s = frame_render_plain(E, F, &st);
    else s = frame_render_batched(E, F, &st);
    if (s != FG_OK) return s;
    F->otf_applied = false;
    if (F->desc.scan.otf) {
        HIP_TRY(launch_otf3(F->d_dneg.as<float>(), F->d_dneg2.as<float>(), F->desc.width, F->desc.height,
                            E->stream), FG_ERR_PIPELINE);
        F->otf_applied = true;
    }
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    st.ms_total = ms_since(t0);
    F->rendered = true;
    F->last_mode = mode;
    if (stats) *stats = st;
    return FG_OK;
}

extern "C" fg_status fg_frame_read(fg_engine *E, const fg_frame *F, float *out)
{
    if (!E || !F || !out) return fail(FG_ERR_INVALID_ARG, "fg_frame_read: null argument");
    if (!F->rendered) return fail(FG_ERR_INVALID_ARG, "fg_frame_read: frame not rendered");
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    const void *src = F->otf_applied ? F->d_dneg2.p : F->d_dneg.p;
    if (F->desc.overlap) {
        // the Hann blend (image_pipeline.rs:593-654) is one of the compositing kernels,
        // which are switched off in this build (device memory fault under investigation);
        // fg_frame_read_tiles returns the tiles
        if (!fg::kFilmBlendEnabled)
            return fail(FG_ERR_PIPELINE, "fg_frame_read: the Hann blend of overlapping tiles is switched off "
                                         "in this build; read the tiles with fg_frame_read_tiles");
        HIP_TRY(fg::launch_film_blend(F->layout, F->d_dneg.as<float>(), F->d_dneg2.as<float>(), E->stream),
                FG_ERR_PIPELINE);
        src = F->d_dneg2.p;
    }
    HIP_TRY(hipMemcpyAsync(out, src, F->emap.size() * 4, hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    return FG_OK;
}

extern "C" fg_status fg_frame_read_tiles(fg_engine *E, const fg_frame *F, int64_t n, float *out,
                                         int64_t *n_total)
{
    if (!E || !F) return fail(FG_ERR_INVALID_ARG, "fg_frame_read_tiles: null argument");
    int64_t total = 0;
    for (const TileRec &t : F->tiles) total += (int64_t)t.tw * t.th;
    if (n_total) *n_total = total;
    if (!out) return FG_OK;
    if (n != total) return fail(FG_ERR_INVALID_ARG, "fg_frame_read_tiles: n must equal the tile values");
    if (!F->rendered) return fail(FG_ERR_INVALID_ARG, "fg_frame_read_tiles: frame not rendered");
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    if (F->desc.overlap) {
        HIP_TRY(hipMemcpy(out, F->d_dneg.p, (size_t)total * 4, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
        return FG_OK;
    }
    // abutting tiles share the frame image: gather each tile's rectangle
    for (const TileRec &t : F->tiles) {
        HIP_TRY(hipMemcpy2D(out, (size_t)t.tw * 4, F->d_dneg.as<float>() + t.out_base, (size_t)F->desc.width * 4,
                            (size_t)t.tw * 4, t.th, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
        out += (size_t)t.tw * t.th;
    }
    return FG_OK;
}

const float *frame_tile_dneg(const fg_frame *F) { return F->d_dneg.as<float>(); }
const fg::FilmLayout *frame_layout(const fg_frame *F) { return F->desc.overlap ? &F->layout : nullptr; }

extern "C" fg_status fg_frame_read_counts(fg_engine *E, const fg_frame *F, int64_t n, uint32_t *counts)
{
    if (!E || !F || !counts || n != F->n_crystals) return fail(FG_ERR_INVALID_ARG, "fg_frame_read_counts: bad argument");
    if (!F->rendered) return fail(FG_ERR_INVALID_ARG, "fg_frame_read_counts: frame not rendered");
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    std::vector<uint32_t> c((size_t)n);
    if (n) HIP_TRY(hipMemcpy(c.data(), F->d_counts.p, (size_t)n * 4, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    for (size_t k = 0; k < F->tiles.size(); k++) {
        const TileRec &t = F->tiles[k];
        const Emulsion &em = F->ems[k];
        for (int64_t j = 0; j < t.n; j++) counts[t.crystal_base + em.bvh.idx[(size_t)j]] = c[(size_t)(t.crystal_base + j)];
    }
    return FG_OK;
}

extern "C" fg_status fg_frame_get_tile(fg_engine *E, const fg_frame *F, int64_t tile, int64_t *n, int64_t *n_nodes,
                                     double *pos, double *dims, double *quat, int32_t *morph, double *sens,
                                     void *nodes, void *records)
{
    if (!E || !F || tile < 0 || tile >= (int64_t)F->tiles.size())
        return fail(FG_ERR_INVALID_ARG, "fg_frame_get_tile: bad argument");
    const TileRec &t = F->tiles[(size_t)tile];
    const Emulsion &em = F->ems[(size_t)tile];
    if (n) *n = t.n;
    if (n_nodes) *n_nodes = t.n_nodes;
    const size_t m = (size_t)t.n;
    if (morph) memcpy(morph, em.morph.data(), m * sizeof(int32_t));
    if (!em.on_device) {
        if (pos) memcpy(pos, em.pos.data(), 3 * m * sizeof(double));
        if (dims) memcpy(dims, em.dims.data(), 3 * m * sizeof(double));
        if (quat) memcpy(quat, em.quat.data(), 4 * m * sizeof(double));
        if (sens) memcpy(sens, em.sens.data(), m * sizeof(double));
    }
    if (!nodes && !records && !em.on_device) return FG_OK;
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    if (em.on_device && m) {
        const size_t b = (size_t)t.crystal_base;
        if (pos) HIP_TRY(hipMemcpy(pos, F->d_pos.as<double>() + 3 * b, 3 * m * 8, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
        if (dims) HIP_TRY(hipMemcpy(dims, F->d_dims.as<double>() + 3 * b, 3 * m * 8, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
        if (quat) HIP_TRY(hipMemcpy(quat, F->d_quat.as<double>() + 4 * b, 4 * m * 8, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
        if (sens) HIP_TRY(hipMemcpy(sens, F->d_sens.as<double>() + b, m * 8, hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    }
    if (nodes && t.n_nodes)
        HIP_TRY(hipMemcpy(nodes, F->d_nodes.as<Node>() + t.node_base, (size_t)t.n_nodes * sizeof(Node),
                          hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    if (records && m)
        HIP_TRY(hipMemcpy(records, F->d_recs.as<Crystal>() + t.crystal_base, m * sizeof(Crystal),
                          hipMemcpyDeviceToHost), FG_ERR_BUFFER_MAP);
    return FG_OK;
}

extern "C" fg_status fg_frame_read_development(fg_engine *E, const fg_frame *F, int64_t n, uint8_t *developed,
                                               double *fe, int64_t *clump_id)
{
    if (!E || !F || n != F->n_crystals) return fail(FG_ERR_INVALID_ARG, "fg_frame_read_development: bad argument");
    if (!F->rendered) return fail(FG_ERR_INVALID_ARG, "fg_frame_read_development: frame not rendered");
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    std::vector<uint8_t> d;
    std::vector<double> f;
    std::vector<int32_t> c;
    const bool plain = F->last_mode == FG_FRAME_PLAIN;
    if (!plain) HIP_TRY(batched_read_develop(F->batched, E->stream, d, f, c), FG_ERR_BUFFER_MAP);
    for (size_t k = 0; k < F->tiles.size(); k++) {
        const TileRec &t = F->tiles[k];
        const Emulsion &em = F->ems[k];
        // slots of the frame's batched pass, or of the tile's own one-tile frame
        if (plain) HIP_TRY(batched_read_develop(F->plain[k], E->stream, d, f, c), FG_ERR_BUFFER_MAP);
        const size_t base = plain ? 0 : (size_t)t.crystal_base;
        for (int64_t j = 0; j < t.n; j++) {
            const size_t sl = base + (size_t)j;
            const size_t o = (size_t)t.crystal_base + em.bvh.idx[(size_t)j];
            if (developed) developed[o] = d[sl];
            if (fe) fe[o] = f[sl];
            if (clump_id) clump_id[o] = c[sl];
        }
    }
    return FG_OK;
}
