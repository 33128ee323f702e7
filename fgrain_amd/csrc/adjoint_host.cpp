// C ABI of the expected (differentiable) model: fg_expected_forward and its
// vector-Jacobian product fg_vjp.  No reference counterpart; the forward restates the
// reference path's expectation (transport.rs:117-134 mean latent count,
// development.rs:69-84 threshold in expectation, transmittance.rs:122-152 scan).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "adjoint.h"
#include "engine_state.h"

using namespace fg;

namespace {

using Clock = std::chrono::steady_clock;

bool valid_model(const fg_expected_model *m, const float *emap, uint32_t tw, uint32_t th,
                 std::string &msg)
{
    if (!m) { msg = "null model"; return false; }
    if (!emap || tw == 0 || th == 0) { msg = "an exposure map is required"; return false; }
    if (!valid_exposure(&m->exposure, msg)) return false;
    if (m->exposure.photons_per_pixel == 0) { msg = "photons_per_pixel must be > 0"; return false; }
    if (!valid_scan(&m->scan, msg)) return false;
    if (m->scan.otf != 0) { msg = "the expected model has no OTF stage (otf must be 0)"; return false; }
    if (!(m->optics.anisotropy_g > -1.0 && m->optics.anisotropy_g < 1.0)) {
        msg = "anisotropy g must lie in (-1, 1)";
        return false;
    }
    if (!(m->optics.antihalation_reflectance >= 0.0)) { msg = "antihalation reflectance must be >= 0"; return false; }
    if (!(m->development.filament_expansion > 0.0)) { msg = "filament expansion must be > 0"; return false; }
    if (!(m->weight_cutoff >= 0.0 && m->weight_cutoff < 1.0)) { msg = "weight cutoff must lie in [0, 1)"; return false; }
    const double packets = (double)m->exposure.photons_per_pixel * tw * th;
    if (packets > 256.0 * 2147483647.0) { msg = "too many packets for one launch"; return false; }
    return true;
}

// Largest expected clump boost (development.rs:164-171 at clump size >= 6).
constexpr double kMaxBoost = 2.05;

// Expected-model scan BVH: every crystal, boxes by the reference's developed-crystal
// rule (crystal.rs:106-133) at the largest boosted expansion fe * kMaxBoost, records with
// semi-axes times fe carrying transport slots; and the touch graph in slot order.
fg_status ensure_expected_bvh(fg_engine *E, Emulsion &em, double fe, double pf)
{
    DeviceEmulsion &d = *em.dev;
    if (d.exp_fe == fe && d.exp_pf == pf) return FG_OK;
    std::vector<uint8_t> all((size_t)em.n, 1u);
    std::vector<double> fes((size_t)em.n, fe), fmax_((size_t)em.n, fe * kMaxBoost);
    RefBvh b;
    build_ref_bvh(em.n, em.pos.data(), em.dims.data(), all.data(), fmax_.data(), nullptr, b);
    std::vector<Node> nodes;
    std::string msg;
    if (!flatten_bvh(b, nodes, msg)) return fail(FG_ERR_INVALID_ARG, "expected scan: " + msg);
    std::vector<Crystal> recs;
    make_records(em, b, all.data(), fes.data(), recs);
    std::vector<uint32_t> slot_of((size_t)em.n);
    for (size_t k = 0; k < em.bvh.idx.size(); k++) slot_of[em.bvh.idx[k]] = (uint32_t)k;
    for (size_t k = 0; k < recs.size(); k++) {
        const uint32_t s = slot_of[b.idx[k]];
        memcpy(&recs[k].sens, &s, sizeof s);
    }
    std::vector<uint32_t> cptr, cidx, sptr((size_t)em.n + 1, 0u), sidx;
    touch_graph(em, fe, pf, cptr, cidx);
    sidx.reserve(cidx.size());
    for (size_t k = 0; k < em.bvh.idx.size(); k++) {
        const uint32_t i = em.bvh.idx[k];
        for (uint32_t e = cptr[i]; e < cptr[i + 1]; e++) sidx.push_back(slot_of[cidx[e]]);
        sptr[k + 1] = (uint32_t)sidx.size();
    }
    HIP_TRY(d.exp_nodes.ensure(nodes.size() * sizeof(Node)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(d.exp_recs.ensure(recs.size() * sizeof(Crystal)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(d.exp_ptr.ensure(sptr.size() * sizeof(uint32_t)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(d.exp_idx.ensure(std::max<size_t>(sidx.size(), 1) * sizeof(uint32_t)), FG_ERR_BUFFER_RANGE);
    if (!nodes.empty())
        HIP_TRY(hipMemcpyAsync(d.exp_nodes.p, nodes.data(), nodes.size() * sizeof(Node),
                               hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    if (!recs.empty())
        HIP_TRY(hipMemcpyAsync(d.exp_recs.p, recs.data(), recs.size() * sizeof(Crystal),
                               hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpyAsync(d.exp_ptr.p, sptr.data(), sptr.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           E->stream), FG_ERR_BUFFER_MAP);
    if (!sidx.empty())
        HIP_TRY(hipMemcpyAsync(d.exp_idx.p, sidx.data(), sidx.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice, E->stream), FG_ERR_BUFFER_MAP);
    // the host vectors go out of scope: wait for the copies
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    d.exp_n_nodes = (uint32_t)nodes.size();
    d.exp_fe = fe;
    d.exp_pf = pf;
    return FG_OK;
}

struct Setup {
    ReplayArgs r{};
    ExpScanArgs s{};
    const uint32_t *ptr = nullptr, *idx = nullptr;  // touch graph, slot order
    uint32_t n = 0;
    size_t npx = 0, nmap = 0;
};

fg_status setup(fg_engine *E, Emulsion &em, const fg_expected_model &m, const float *emap,
                uint32_t tw, uint32_t th, Setup &S)
{
    fg_status st = ensure_device(E, em);
    if (st != FG_OK) return st;
    st = ensure_expected_bvh(E, em, m.development.filament_expansion,
                             m.development.clumping_proximity_factor);
    if (st != FG_OK) return st;
    DeviceEmulsion &d = *em.dev;
    S.n = (uint32_t)em.n;
    S.nmap = (size_t)tw * th;
    S.npx = (size_t)m.scan.width_px * m.scan.height_px;
    fg_emulsion_config ec = em.cfg;
    ec.optics = m.optics;
    TransportArgs &a = S.r.t;
    a.nodes = d.nodes.as<Node>();
    a.recs = d.recs.as<float4>();
    a.pre = d.pre.as<PreRec>();
    a.n_nodes = d.n_nodes;
    fill_transport(a, ec, m.exposure, tw, th, m.expose_seed, em.reach, E->tuning);
    HIP_TRY(E->emap.ensure(S.nmap * sizeof(float)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemcpyAsync(E->emap.p, emap, S.nmap * sizeof(float), hipMemcpyHostToDevice, E->stream),
            FG_ERR_BUFFER_MAP);
    a.emap = E->emap.as<float>();
    a.estride = tw;
    HIP_TRY(E->stats.ensure(8 * sizeof(unsigned long long)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemsetAsync(E->stats.p, 0, 8 * sizeof(unsigned long long), E->stream), FG_ERR_BUFFER_MAP);
    S.r.stats = E->stats.as<unsigned long long>();
    S.r.hal_unit = (float)pow(10.0, -2.0 * m.optics.antihalation_density);
    S.r.refl = (float)m.optics.antihalation_reflectance;
    S.r.cut = (float)(m.weight_cutoff > 0.0 ? m.weight_cutoff : 1e-6);
    const size_t n = std::max<size_t>(S.n, 1);
    for (DevBuf *b : {&E->dose32, &E->prob, &E->dprob, &E->gdose, &E->boost, &E->dboost})
        HIP_TRY(b->ensure(n * sizeof(float)), FG_ERR_BUFFER_RANGE);
    for (DevBuf *b : {&E->dose64, &E->gprob, &E->gboost})
        HIP_TRY(b->ensure(n * sizeof(double)), FG_ERR_BUFFER_RANGE);
    S.ptr = d.exp_ptr.as<uint32_t>();
    S.idx = d.exp_idx.as<uint32_t>();
    ExpScanArgs &s = S.s;
    s.nodes = d.exp_nodes.as<Node>();
    s.recs = d.exp_recs.as<float4>();
    s.n_nodes = d.exp_n_nodes;
    s.w_px = m.scan.width_px;
    s.h_px = m.scan.height_px;
    const uint32_t spp = std::max(m.scan.samples_per_pixel, 1u);
    s.g = std::max((uint32_t)llround(sqrt((double)spp)), 1u);
    s.ss = std::max(m.scan.supersample, 1u);
    s.mode = m.scan.mode;
    s.pw = (float)(em.cfg.width / (double)std::max(m.scan.width_px, 1u));
    s.ph = (float)(em.cfg.height / (double)std::max(m.scan.height_px, 1u));
    s.z0 = (float)(em.cfg.depth + 0.1);
    s.sigma = (float)m.scan.silver_opacity;
    s.key0 = (uint32_t)m.scan_seed;
    s.key1 = (uint32_t)(m.scan_seed >> 32);
    s.rays_per_lane = E->tuning.scan_rays_per_thread;
    s.prob = E->prob.as<float>();
    s.boost = E->boost.as<float>();
    return FG_OK;
}

// The replay kernel the tuning selects (persistent lanes unless one thread per packet
// is asked for); the 64-bit work counter is stats slot 6, zeroed before each pass.
bool replay_persistent(fg_engine *E)
{
    if (E->tuning.transport_kernel == 1) return false;
    if (E->replay_grid == 0) E->replay_grid = E->cus * replay_persistent_blocks_per_cu();
    return true;
}

// expected dose (slot order) into dose32, then prob and dprob
fg_status dose_pass(fg_engine *E, Setup &S, const fg_expected_model &m, float *ms)
{
    if (S.n > 0) HIP_TRY(hipMemsetAsync(E->dose64.p, 0, (size_t)S.n * sizeof(double), E->stream), FG_ERR_BUFFER_MAP);
    S.r.dose = E->dose64.as<double>();
    HIP_TRY(hipEventRecord(E->ev0, E->stream), FG_ERR_CHANNEL);
    if (replay_persistent(E))
        HIP_TRY(launch_replay_persistent_dose(S.r, E->stats.as<unsigned long long>() + 6, E->replay_grid, E->stream),
                FG_ERR_PIPELINE);
    else
        HIP_TRY(launch_replay_dose(S.r, E->stream), FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev1, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(launch_dev_prob(E->dose64.as<double>(), E->dose32.as<float>(), E->prob.as<float>(),
                            E->dprob.as<float>(), S.n, m.development.threshold_photons, E->stream),
            FG_ERR_PIPELINE);
    HIP_TRY(launch_boost(S.ptr, S.idx, E->prob.as<float>(), E->boost.as<float>(), E->dboost.as<float>(), S.n,
                         E->stream), FG_ERR_PIPELINE);
    HIP_TRY(hipEventSynchronize(E->ev1), FG_ERR_CHANNEL);
    if (ms) (void)hipEventElapsedTime(ms, E->ev0, E->ev1);
    return FG_OK;
}

void to_crystal_order(const Emulsion &em, const std::vector<float> &slot_vals, float *out)
{
    for (size_t k = 0; k < slot_vals.size(); k++) out[em.bvh.idx[k]] = slot_vals[k];
}

void fill_stats(fg_engine *E, const Setup &S, fg_adjoint_stats *st, float ms_replay, float ms_scan,
                Clock::time_point t0)
{
    if (!st) return;
    unsigned long long c[8] = {0};
    (void)hipMemcpy(c, E->stats.p, sizeof c, hipMemcpyDeviceToHost);
    memset(st, 0, sizeof *st);
    st->packets = S.r.t.n_photons;
    st->bounces = c[1];
    st->events = c[2];
    st->atomics = c[3];
    st->ms_replay = ms_replay;
    st->ms_scan = ms_scan;
    st->ms_total = std::chrono::duration<float, std::milli>(Clock::now() - t0).count();
}

}  // namespace

extern "C" fg_status fg_expected_forward(fg_engine *E, const fg_emulsion *emw,
                                         const fg_expected_model *model, const float *emap,
                                         uint32_t tw, uint32_t th, float *dose, float *prob,
                                         float *out, fg_adjoint_stats *stats)
{
    const auto t0 = Clock::now();
    std::string msg;
    if (!E || !emw || !out) return fail(FG_ERR_INVALID_ARG, "fg_expected_forward: null argument");
    if (!valid_model(model, emap, tw, th, msg)) return fail(FG_ERR_INVALID_ARG, "fg_expected_forward: " + msg);
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    Emulsion &em = const_cast<fg_emulsion *>(emw)->em;
    Setup S;
    fg_status st = setup(E, em, *model, emap, tw, th, S);
    if (st != FG_OK) return st;
    float ms_replay = 0.f, ms_scan = 0.f;
    st = dose_pass(E, S, *model, &ms_replay);
    if (st != FG_OK) return st;
    HIP_TRY(E->out.ensure(std::max<size_t>(S.npx, 1) * sizeof(float)), FG_ERR_BUFFER_RANGE);
    S.s.out = E->out.as<float>();
    HIP_TRY(hipEventRecord(E->ev2, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(launch_expected_scan(S.s, E->stream), FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev3, E->stream), FG_ERR_CHANNEL);
    if (S.npx)
        HIP_TRY(hipMemcpyAsync(out, E->out.p, S.npx * sizeof(float), hipMemcpyDeviceToHost, E->stream),
                FG_ERR_BUFFER_MAP);
    std::vector<float> d32(S.n), p32(S.n);
    if (S.n && dose)
        HIP_TRY(hipMemcpyAsync(d32.data(), E->dose32.p, (size_t)S.n * 4, hipMemcpyDeviceToHost, E->stream),
                FG_ERR_BUFFER_MAP);
    if (S.n && prob)
        HIP_TRY(hipMemcpyAsync(p32.data(), E->prob.p, (size_t)S.n * 4, hipMemcpyDeviceToHost, E->stream),
                FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    (void)hipEventElapsedTime(&ms_scan, E->ev2, E->ev3);
    if (dose) to_crystal_order(em, d32, dose);
    if (prob) to_crystal_order(em, p32, prob);
    fill_stats(E, S, stats, ms_replay, ms_scan, t0);
    return FG_OK;
}

extern "C" fg_status fg_vjp(fg_engine *E, const fg_emulsion *emw, const fg_expected_model *model,
                            const float *emap, uint32_t tw, uint32_t th, const float *dose,
                            const float *grad_out, float *grad_exposure, fg_param_grads *grads,
                            float *grad_dose, fg_adjoint_stats *stats)
{
    const auto t0 = Clock::now();
    std::string msg;
    if (!E || !emw || !grad_out || !grad_exposure || !grads)
        return fail(FG_ERR_INVALID_ARG, "fg_vjp: null argument");
    if (!valid_model(model, emap, tw, th, msg)) return fail(FG_ERR_INVALID_ARG, "fg_vjp: " + msg);
    std::lock_guard<std::mutex> lk(E->mu);
    HIP_TRY(hipSetDevice(E->device), FG_ERR_DEVICE_REQUEST);
    Emulsion &em = const_cast<fg_emulsion *>(emw)->em;
    Setup S;
    fg_status st = setup(E, em, *model, emap, tw, th, S);
    if (st != FG_OK) return st;
    float ms_replay = 0.f, ms_scan = 0.f;
    if (dose) {
        std::vector<float> slot((size_t)S.n);
        for (size_t k = 0; k < slot.size(); k++) slot[k] = dose[em.bvh.idx[k]];
        if (S.n)
            HIP_TRY(hipMemcpyAsync(E->dose32.p, slot.data(), (size_t)S.n * 4, hipMemcpyHostToDevice, E->stream),
                    FG_ERR_BUFFER_MAP);
        HIP_TRY(launch_dev_prob_from(E->dose32.as<float>(), E->prob.as<float>(), E->dprob.as<float>(), S.n,
                                     model->development.threshold_photons, E->stream),
                FG_ERR_PIPELINE);
        HIP_TRY(launch_boost(S.ptr, S.idx, E->prob.as<float>(), E->boost.as<float>(), E->dboost.as<float>(),
                             S.n, E->stream), FG_ERR_PIPELINE);
        HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    } else {
        st = dose_pass(E, S, *model, nullptr);
        if (st != FG_OK) return st;
        HIP_TRY(hipMemsetAsync(E->stats.p, 0, 8 * sizeof(unsigned long long), E->stream), FG_ERR_BUFFER_MAP);
    }
    // scan adjoint: dL/dprob, then dL/ddose
    HIP_TRY(E->gout.ensure(std::max<size_t>(S.npx, 1) * sizeof(float)), FG_ERR_BUFFER_RANGE);
    if (S.npx)
        HIP_TRY(hipMemcpyAsync(E->gout.p, grad_out, S.npx * sizeof(float), hipMemcpyHostToDevice, E->stream),
                FG_ERR_BUFFER_MAP);
    if (S.n) HIP_TRY(hipMemsetAsync(E->gprob.p, 0, (size_t)S.n * sizeof(double), E->stream), FG_ERR_BUFFER_MAP);
    if (S.n) HIP_TRY(hipMemsetAsync(E->gboost.p, 0, (size_t)S.n * sizeof(double), E->stream), FG_ERR_BUFFER_MAP);
    S.s.gout = E->gout.as<float>();
    S.s.gprob = E->gprob.as<double>();
    S.s.gboost = E->gboost.as<double>();
    HIP_TRY(hipEventRecord(E->ev2, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(launch_expected_scan_adjoint(S.s, E->stream), FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev3, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(launch_boost_grad(S.ptr, S.idx, E->gboost.as<double>(), E->dboost.as<float>(), E->gprob.as<double>(),
                              S.n, E->stream), FG_ERR_PIPELINE);
    HIP_TRY(launch_dose_grad(E->gprob.as<double>(), E->dprob.as<float>(), E->gdose.as<float>(), S.n, E->stream),
            FG_ERR_PIPELINE);
    // gradient replay
    const bool persistent = replay_persistent(E);
    const uint32_t blocks = persistent ? (uint32_t)E->replay_grid : (uint32_t)replay_blocks(S.r.t.n_photons);
    HIP_TRY(E->gmap.ensure(S.nmap * sizeof(double)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->partials.ensure(std::max<size_t>(blocks, 1) * 4 * sizeof(double)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->pgrad.ensure(4 * sizeof(double)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(E->out2.ensure(S.nmap * sizeof(float)), FG_ERR_BUFFER_RANGE);
    HIP_TRY(hipMemsetAsync(E->gmap.p, 0, S.nmap * sizeof(double), E->stream), FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemsetAsync(E->partials.p, 0, std::max<size_t>(blocks, 1) * 4 * sizeof(double), E->stream),
            FG_ERR_BUFFER_MAP);
    S.r.gdose = E->gdose.as<float>();
    S.r.gmap = E->gmap.as<double>();
    S.r.partials = E->partials.as<double>();
    HIP_TRY(hipEventRecord(E->ev0, E->stream), FG_ERR_CHANNEL);
    if (persistent)
        HIP_TRY(launch_replay_persistent_grad(S.r, E->stats.as<unsigned long long>() + 6, E->replay_grid, E->stream),
                FG_ERR_PIPELINE);
    else
        HIP_TRY(launch_replay_grad(S.r, E->stream), FG_ERR_PIPELINE);
    HIP_TRY(hipEventRecord(E->ev1, E->stream), FG_ERR_CHANNEL);
    HIP_TRY(launch_reduce_partials(E->partials.as<double>(), blocks, E->pgrad.as<double>(), E->stream),
            FG_ERR_PIPELINE);
    HIP_TRY(launch_f64_to_f32(E->gmap.as<double>(), E->out2.as<float>(), (uint32_t)S.nmap, E->stream),
            FG_ERR_PIPELINE);
    double pg[4];
    HIP_TRY(hipMemcpyAsync(grad_exposure, E->out2.p, S.nmap * sizeof(float), hipMemcpyDeviceToHost, E->stream),
            FG_ERR_BUFFER_MAP);
    HIP_TRY(hipMemcpyAsync(pg, E->pgrad.p, sizeof pg, hipMemcpyDeviceToHost, E->stream), FG_ERR_BUFFER_MAP);
    std::vector<float> gd(S.n);
    if (S.n && grad_dose)
        HIP_TRY(hipMemcpyAsync(gd.data(), E->gdose.p, (size_t)S.n * 4, hipMemcpyDeviceToHost, E->stream),
                FG_ERR_BUFFER_MAP);
    HIP_TRY(hipStreamSynchronize(E->stream), FG_ERR_CHANNEL);
    (void)hipEventElapsedTime(&ms_replay, E->ev0, E->ev1);
    (void)hipEventElapsedTime(&ms_scan, E->ev2, E->ev3);
    grads->crystal_absorption_coeff = pg[0];
    grads->anisotropy_g = pg[1];
    grads->antihalation_reflectance = pg[2];
    grads->reserved_ = 0.0;
    if (grad_dose) to_crystal_order(em, gd, grad_dose);
    fill_stats(E, S, stats, ms_replay, ms_scan, t0);
    return FG_OK;
}
