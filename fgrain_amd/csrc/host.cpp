// Host-side emulsion work: synthesis, the reference-exact BVH build, the threaded
// device layout and host development.  Compiled with -ffp-contract=off.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <numeric>
#include <unordered_map>

#include "internal.h"

namespace fg {

static constexpr double kTau = 6.283185307179586;
static constexpr double kPi = 3.141592653589793;

uint32_t SeqRng::u32()
{
    if (pos == 4) {
        buf[0] = (uint32_t)block;
        buf[1] = (uint32_t)(block >> 32);
        buf[2] = 0u;
        buf[3] = kTagSeq;
        philox(buf, key[0], key[1]);
        block++;
        pos = 0;
    }
    return buf[pos++];
}

// Box-Muller; one normal from two uniforms.
double SeqRng::normal(double mean, double sd)
{
    const double u1 = 1.0 - uniform(), u2 = uniform();
    return mean + sd * (sqrt(-2.0 * log(u1)) * cos(kTau * u2));
}

// sample_poisson_3d (poisson.rs:14-132): serial Bridson with a grid of cell r/sqrt(3),
// k candidates per active point in the shell [r, 2r] (radius cbrt(7u+1) r), neighbour
// check over +-2 cells, swap_remove on failure.
static void poisson3d(double W, double H, double D, double r, int k, SeqRng &rng,
                      std::vector<double> &pts)
{
    pts.clear();
    if (W <= 0 || H <= 0 || D <= 0 || r <= 0) return;
    const double cell = r / sqrt(3.0);
    const int64_t gw = (int64_t)ceil(W / cell), gh = (int64_t)ceil(H / cell),
                  gd = (int64_t)ceil(D / cell);
    std::vector<int64_t> grid((size_t)(gw * gh * gd), -1);
    std::vector<int64_t> active;
    auto cell_of = [&](const double *p, int64_t &x, int64_t &y, int64_t &z) {
        x = std::min(std::max((int64_t)floor(p[0] / cell), (int64_t)0), gw - 1);
        y = std::min(std::max((int64_t)floor(p[1] / cell), (int64_t)0), gh - 1);
        z = std::min(std::max((int64_t)floor(p[2] / cell), (int64_t)0), gd - 1);
    };
    {
        const double x = rng.uniform() * W;
        const double y = rng.uniform() * H;
        const double z = rng.uniform() * D;
        pts.push_back(x);
        pts.push_back(y);
        pts.push_back(z);
        int64_t gx, gy, gz;
        cell_of(pts.data(), gx, gy, gz);
        grid[(size_t)(gx + gy * gw + gz * gw * gh)] = 0;
        active.push_back(0);
    }
    const double r2 = r * r;
    while (!active.empty()) {
        const size_t pick = (size_t)rng.below(active.size());
        const double bx = pts[3 * active[pick]], by = pts[3 * active[pick] + 1],
                     bz = pts[3 * active[pick] + 2];
        bool found = false;
        for (int s = 0; s < k; s++) {
            const double u = rng.uniform();
            const double ct = -1.0 + 2.0 * rng.uniform();
            const double st = sqrt(std::max(1.0 - ct * ct, 0.0));
            const double phi = rng.uniform() * kTau;
            const double rad = cbrt(7.0 * u + 1.0) * r;
            const double c[3] = {bx + rad * st * cos(phi), by + rad * st * sin(phi), bz + rad * ct};
            if (c[0] < 0.0 || c[0] >= W || c[1] < 0.0 || c[1] >= H || c[2] < 0.0 || c[2] >= D)
                continue;
            int64_t cx, cy, cz;
            cell_of(c, cx, cy, cz);
            bool close = false;
            for (int64_t z = std::max(cz - 2, (int64_t)0); z <= std::min(cz + 2, gd - 1) && !close; z++)
                for (int64_t y = std::max(cy - 2, (int64_t)0); y <= std::min(cy + 2, gh - 1) && !close; y++)
                    for (int64_t x = std::max(cx - 2, (int64_t)0); x <= std::min(cx + 2, gw - 1); x++) {
                        const int64_t e = grid[(size_t)(x + y * gw + z * gw * gh)];
                        if (e < 0) continue;
                        const double dx = pts[3 * e] - c[0], dy = pts[3 * e + 1] - c[1],
                                     dz = pts[3 * e + 2] - c[2];
                        if (dx * dx + dy * dy + dz * dz < r2) {
                            close = true;
                            break;
                        }
                    }
            if (!close) {
                const int64_t id = (int64_t)(pts.size() / 3);
                grid[(size_t)(cx + cy * gw + cz * gw * gh)] = id;
                pts.insert(pts.end(), c, c + 3);
                active.push_back(id);
                found = true;
                break;
            }
        }
        if (!found) {
            active[pick] = active.back();
            active.pop_back();
        }
    }
}

// Per-point Philox stream of device synthesis: ctr = (lo32(p), hi32(p), block, kTagAttr).
struct PointRng {
    uint32_t key[2], p0, p1;
    uint32_t block = 0, buf[4];
    int pos = 4;
    PointRng(uint64_t seed, uint64_t p)
    {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
        p0 = (uint32_t)p;
        p1 = (uint32_t)(p >> 32);
    }
    uint32_t u32()
    {
        if (pos == 4) {
            buf[0] = p0;
            buf[1] = p1;
            buf[2] = block++;
            buf[3] = kTagAttr;
            philox(buf, key[0], key[1]);
            pos = 0;
        }
        return buf[pos++];
    }
    double uniform()
    {
        const uint64_t lo = u32(), hi = u32();
        return (double)((lo | (hi << 32)) >> 11) * 0x1p-53;
    }
    double normal(double mean, double sd)
    {
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return mean + sd * (sqrt(-2.0 * log(u1)) * cos(kTau * u2));
    }
};

// Crystal attributes of emulsion.rs:160-207 for every point of em.pos; rng_of(i) gives
// the stream crystal i draws from.
template <class RngOf>
static void crystal_attributes(const fg_emulsion_config &cfg, Emulsion &em, RngOf &&rng_of)
{
    const int64_t n = em.n;
    em.dims.assign((size_t)(3 * n), 0.0);
    em.quat.assign((size_t)(4 * n), 0.0);
    em.sens.assign((size_t)n, 0.0);
    em.morph.assign((size_t)n, (int32_t)cfg.morphology);
    const double mu = log(cfg.mean_crystal_radius);
    const double sd_sens = cfg.sensitivity_sigma;
    for (int64_t i = 0; i < n; i++) {
        auto &rng = rng_of(i);
        double rad = exp(rng.normal(mu, cfg.size_variance));
        rad = std::min(std::max(rad, cfg.min_radius), cfg.r_min * 0.48);
        const double s = std::max(rng.normal(cfg.base_sensitivity, sd_sens), 0.1);
        double *q = &em.quat[(size_t)(4 * i)];
        if (cfg.morphology == 0 || cfg.morphology == 2) {
            const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
            // UnitQuaternion::from_euler_angles(2 pi u1, pi u2, 2 pi u3)
            const double ro = u1 * kTau * 0.5, pi_ = u2 * kPi * 0.5, ya = u3 * kTau * 0.5;
            const double sr = sin(ro), cr = cos(ro), sp = sin(pi_), cp = cos(pi_), sy = sin(ya),
                         cy = cos(ya);
            q[0] = cr * cp * cy + sr * sp * sy;
            q[1] = sr * cp * cy - cr * sp * sy;
            q[2] = cr * sp * cy + sr * cp * sy;
            q[3] = cr * cp * sy - sr * sp * cy;
            em.dims[(size_t)(3 * i)] = em.dims[(size_t)(3 * i + 1)] = em.dims[(size_t)(3 * i + 2)] = rad;
        } else {
            const double ht = std::max(rad / cfg.tabular_aspect_ratio, cfg.min_half_thickness);
            const double tilt = -0.17 + 0.34 * rng.uniform();
            double ax[3] = {-1.0 + 2.0 * rng.uniform(), -1.0 + 2.0 * rng.uniform(), 0.0};
            const double nn = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
            ax[0] /= nn;
            ax[1] /= nn;
            ax[2] /= nn;
            const double sh = sin(tilt * 0.5), ch = cos(tilt * 0.5);
            q[0] = ch;
            q[1] = ax[0] * sh;
            q[2] = ax[1] * sh;
            q[3] = ax[2] * sh;
            em.dims[(size_t)(3 * i)] = rad;
            em.dims[(size_t)(3 * i + 1)] = rad;
            em.dims[(size_t)(3 * i + 2)] = ht;
        }
        em.sens[(size_t)i] = s;
    }
}

// VirtualEmulsionLayer::synthesize (emulsion.rs:155-211)
void synthesize(const fg_emulsion_config &cfg, uint64_t seed, Emulsion &em)
{
    SeqRng rng(seed);
    std::vector<double> pts;
    const int k = cfg.k_samples ? (int)cfg.k_samples : 30;
    poisson3d(cfg.width, cfg.height, cfg.depth, cfg.r_min, k, rng, pts);
    em.cfg = cfg;
    em.n = (int64_t)(pts.size() / 3);
    em.pos = pts;
    crystal_attributes(cfg, em, [&](int64_t) -> SeqRng & { return rng; });
}

void attributes_per_point(const fg_emulsion_config &cfg, uint64_t seed, Emulsion &em)
{
    PointRng rng(seed, 0);
    crystal_attributes(cfg, em, [&](int64_t i) -> PointRng & {
        rng = PointRng(seed, (uint64_t)i);
        return rng;
    });
}

// ------------------------------------------------------------------------ BVH ---
namespace {
struct Builder {
    RefBvh *b;
    const double *amn, *amx, *cen;

    // BVH::subdivide (bvh.rs:166-295): leaf at <= 4 primitives or a degenerate
    // centroid extent; median of the centroid box on its longest axis (ties x > y > z);
    // in-place two-pointer partition; count/2 fallback; interior reserved before its
    // children so the left child is node + 1.
    uint32_t subdivide(int64_t start, int64_t count)
    {
        const uint32_t node = (uint32_t)b->left.size();
        double bmn[3] = {INFINITY, INFINITY, INFINITY}, bmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        double cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t k = start; k < start + count; k++) {
            const uint32_t id = b->idx[(size_t)k];
            for (int a = 0; a < 3; a++) {
                bmn[a] = fmin(bmn[a], amn[3 * id + a]);
                bmx[a] = fmax(bmx[a], amx[3 * id + a]);
                cmn[a] = fmin(cmn[a], cen[3 * id + a]);
                cmx[a] = fmax(cmx[a], cen[3 * id + a]);
            }
        }
        b->mn.insert(b->mn.end(), bmn, bmn + 3);
        b->mx.insert(b->mx.end(), bmx, bmx + 3);
        b->left.push_back(0);
        b->right.push_back(0);
        b->offset.push_back(0);
        b->count.push_back(0);
        auto leaf = [&]() {
            b->offset[node] = (uint32_t)start;
            b->count[node] = (uint32_t)count;
            return node;
        };
        if (count <= 4) return leaf();
        const double ext[3] = {cmx[0] - cmn[0], cmx[1] - cmn[1], cmx[2] - cmn[2]};
        const int axis = (ext[0] >= ext[1] && ext[0] >= ext[2]) ? 0 : (ext[1] >= ext[2] ? 1 : 2);
        if (ext[axis] <= 1e-7) return leaf();
        const double mid = 0.5 * (cmn[axis] + cmx[axis]);
        int64_t i = start, j = start + count - 1;
        while (i <= j) {
            if (cen[3 * b->idx[(size_t)i] + axis] < mid) {
                i++;
            } else {
                std::swap(b->idx[(size_t)i], b->idx[(size_t)j]);
                if (j == 0) break;
                j--;
            }
        }
        const int64_t lc = i - start;
        const int64_t split = (lc == 0 || lc == count) ? count / 2 : lc;
        const uint32_t l = subdivide(start, split);
        const uint32_t r = subdivide(start + split, count - split);
        b->left[node] = l;
        b->right[node] = r;
        return node;
    }
};
}  // namespace

// BVH::build (bvh.rs:157-195).  subset (optional) lists the crystals to include; the
// BVH's indices then refer to positions in subset, as when the reference builds over a
// cloned sub-vector (transmittance.rs:82-89).  Boxes follow HalideCrystal::aabb
// (crystal.rs:106-133) with expansion = developed ? fe : 1.
void build_ref_bvh(int64_t n, const double *pos, const double *dims, const uint8_t *developed,
                   const double *fe, const uint32_t *subset, RefBvh &out)
{
    out = RefBvh();
    if (n == 0) return;
    std::vector<double> amn((size_t)(3 * n)), amx((size_t)(3 * n)), cen((size_t)(3 * n));
    out.idx.resize((size_t)n);
    for (int64_t k = 0; k < n; k++) {
        const int64_t i = subset ? (int64_t)subset[k] : k;
        const double e = (developed && developed[i]) ? (fe ? fe[i] : 1.0) : 1.0;
        const double m = fmax(fmax(dims[3 * i], dims[3 * i + 1]), dims[3 * i + 2]) * e;
        for (int a = 0; a < 3; a++) {
            amn[(size_t)(3 * k + a)] = pos[3 * i + a] - m;
            amx[(size_t)(3 * k + a)] = pos[3 * i + a] + m;
            cen[(size_t)(3 * k + a)] = 0.5 * (amn[(size_t)(3 * k + a)] + amx[(size_t)(3 * k + a)]);
        }
        out.idx[(size_t)k] = (uint32_t)k;
    }
    out.mn.reserve((size_t)(6 * n));
    out.mx.reserve((size_t)(6 * n));
    Builder B{&out, amn.data(), amx.data(), cen.data()};
    B.subdivide(0, n);
}

float round_down_f(double x)
{
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}

float round_up_f(double x)
{
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}

bool flatten_bvh(const RefBvh &b, std::vector<Node> &out, std::string &msg)
{
    out.clear();
    const int64_t nn = b.nodes();
    if (nn == 0) return true;
    std::vector<uint32_t> size((size_t)nn, 1);
    std::vector<int> depth((size_t)nn, 0);
    // children always follow their parent in the reference layout: sizes bottom-up
    for (int64_t i = nn - 1; i >= 0; i--)
        if (b.count[(size_t)i] == 0) size[(size_t)i] = 1 + size[b.left[(size_t)i]] + size[b.right[(size_t)i]];
    int max_interior_depth = 0;
    for (int64_t i = 0; i < nn; i++) {
        if (b.count[(size_t)i] == 0) {
            depth[b.left[(size_t)i]] = depth[(size_t)i] + 1;
            depth[b.right[(size_t)i]] = depth[(size_t)i] + 1;
            max_interior_depth = std::max(max_interior_depth, depth[(size_t)i]);
        } else if (b.count[(size_t)i] > kMaxLeafCount) {
            msg = "BVH leaf holds more than 15 crystals (degenerate centroids)";
            return false;
        }
    }
    // The reference drops both children of a node popped with 62 or more pending
    // entries (bvh.rs:339); pending entries never exceed the node depth, so for
    // interior depths <= 61 the stackless walk is identical.
    if (max_interior_depth > 61) {
        msg = "BVH deeper than the reference traversal stack";
        return false;
    }
    out.resize((size_t)nn);
    // iterative emission: (ref node, slot, skip)
    struct Item { uint32_t r, slot, skip; };
    std::vector<Item> st;
    st.push_back({0u, 0u, (uint32_t)nn});
    while (!st.empty()) {
        const Item it = st.back();
        st.pop_back();
        Node &o = out[it.slot];
        for (int a = 0; a < 3; a++) {
            o.lo[a] = round_down_f(b.mn[3 * (size_t)it.r + a]);
            o.hi[a] = round_up_f(b.mx[3 * (size_t)it.r + a]);
        }
        o.skip = it.skip;
        if (b.count[it.r] > 0) {
            o.prim = (b.offset[it.r] << 4) | b.count[it.r];
        } else {
            o.prim = kInterior;
            const uint32_t rs = it.slot + 1;
            const uint32_t ls = rs + size[b.right[it.r]];
            st.push_back({b.right[it.r], rs, ls});
            st.push_back({b.left[it.r], ls, it.skip});
        }
    }
    return true;
}

void make_records(const Emulsion &em, const RefBvh &b, const uint8_t *developed,
                  const double *fe, std::vector<Crystal> &out)
{
    out.resize(b.idx.size());
    for (size_t k = 0; k < b.idx.size(); k++) {
        const int64_t i = b.idx[k];
        const double e = (developed && developed[i]) ? (fe ? fe[i] : 1.0) : 1.0;
        Crystal &r = out[k];
        for (int a = 0; a < 3; a++) {
            r.c[a] = (float)em.pos[(size_t)(3 * i + a)];
            r.h[a] = (float)(em.dims[(size_t)(3 * i + a)] * e);
        }
        for (int a = 0; a < 4; a++) r.q[a] = (float)em.quat[(size_t)(4 * i + a)];
        r.sens = (float)em.sens[(size_t)i];
        r.morph = (uint32_t)em.morph[(size_t)i];
    }
}

bool finalize_emulsion(Emulsion &em, std::string &msg)
{
    em.counts.assign((size_t)em.n, 0u);
    em.developed.assign((size_t)em.n, 0u);
    em.fe.assign((size_t)em.n, 1.0);
    em.clump.assign((size_t)em.n, -1);
    build_ref_bvh(em.n, em.pos.data(), em.dims.data(), nullptr, nullptr, nullptr, em.bvh);
    if (!flatten_bvh(em.bvh, em.nodes, msg)) return false;
    make_records(em, em.bvh, nullptr, nullptr, em.recs);
    em.dev_state_dirty = true;
    double m = 0.0;
    for (int64_t i = 0; i < em.n; i++)
        if (em.morph[(size_t)i] == 0)
            m = std::max(m, fmax(fmax(em.dims[(size_t)(3 * i)], em.dims[(size_t)(3 * i + 1)]),
                                 em.dims[(size_t)(3 * i + 2)]));
    em.reach = (sqrt(3.0) - 1.0) * m;
    return true;
}

// ---------------------------------------------------------------- development ---
// develop_emulsion (development.rs:59-188).  With a uniform filament expansion before
// any boost, the BFS's touch relation |p_i - p_j| <= (m_i fe + m_j fe) pf is
// symmetric and each BFS tree is a connected component, so clumps are computed here
// by union-find over a uniform grid (cell = largest touch distance) and numbered by
// their smallest crystal index, the order the reference discovers them in.
void develop_host(Emulsion &em, const fg_development_config &cfg, fg_development_report &rep)
{
    const int64_t n = em.n;
    memset(&rep, 0, sizeof rep);
    rep.total_crystals = (uint64_t)n;
    if (n == 0) return;
    int64_t ndev = 0;
    for (int64_t i = 0; i < n; i++) {
        if (em.counts[(size_t)i] >= cfg.threshold_photons) {
            em.developed[(size_t)i] = 1;
            em.fe[(size_t)i] = cfg.filament_expansion;
            ndev++;
        } else {
            em.developed[(size_t)i] = 0;
            em.fe[(size_t)i] = 0.0;
        }
        em.clump[(size_t)i] = -1;
    }
    rep.developed_crystals = (uint64_t)ndev;
    rep.fixed_crystals = (uint64_t)(n - ndev);
    std::vector<double> mdim((size_t)n);
    double maxb = 0.0;
    for (int64_t i = 0; i < n; i++) {
        const double *d = &em.dims[(size_t)(3 * i)];
        mdim[(size_t)i] = fmax(fmax(d[0], d[1]), d[2]);
        maxb = fmax(maxb, mdim[(size_t)i]);
    }
    const double fe = cfg.filament_expansion, pf = cfg.clumping_proximity_factor;
    std::vector<int64_t> parent((size_t)n);
    std::iota(parent.begin(), parent.end(), 0);
    std::function<int64_t(int64_t)> find = [&](int64_t x) {
        while (parent[(size_t)x] != x) {
            parent[(size_t)x] = parent[(size_t)parent[(size_t)x]];
            x = parent[(size_t)x];
        }
        return x;
    };
    const double cell = std::max((2.0 * maxb * fe) * pf, 1e-9);
    std::unordered_map<int64_t, std::vector<int64_t>> grid;
    auto key = [](int64_t x, int64_t y, int64_t z) {
        return (x + (1 << 20)) | ((y + (1 << 20)) << 21) | ((z + (1 << 20)) << 42);
    };
    for (int64_t i = 0; i < n; i++) {
        if (!em.developed[(size_t)i]) continue;
        const double *p = &em.pos[(size_t)(3 * i)];
        grid[key((int64_t)floor(p[0] / cell), (int64_t)floor(p[1] / cell), (int64_t)floor(p[2] / cell))].push_back(i);
    }
    for (int64_t i = 0; i < n; i++) {
        if (!em.developed[(size_t)i]) continue;
        const double *p = &em.pos[(size_t)(3 * i)];
        const int64_t gx = (int64_t)floor(p[0] / cell), gy = (int64_t)floor(p[1] / cell),
                      gz = (int64_t)floor(p[2] / cell);
        const double ri = mdim[(size_t)i] * fe;
        for (int64_t z = gz - 1; z <= gz + 1; z++)
            for (int64_t y = gy - 1; y <= gy + 1; y++)
                for (int64_t x = gx - 1; x <= gx + 1; x++) {
                    auto itc = grid.find(key(x, y, z));
                    if (itc == grid.end()) continue;
                    for (int64_t j : itc->second) {
                        if (j <= i) continue;
                        const double rj = mdim[(size_t)j] * fe;
                        const double touch = (ri + rj) * pf;
                        const double *q = &em.pos[(size_t)(3 * j)];
                        const double dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
                        if (dx * dx + dy * dy + dz * dz <= touch * touch) {
                            const int64_t a = find(i), b = find(j);
                            if (a != b) parent[(size_t)std::max(a, b)] = std::min(a, b);
                        }
                    }
                }
    }
    std::vector<int64_t> size((size_t)n, 0), id((size_t)n, -1);
    int64_t clumps = 0;
    for (int64_t i = 0; i < n; i++) {
        if (!em.developed[(size_t)i]) continue;
        const int64_t r = find(i);
        if (r == i) id[(size_t)i] = clumps++;
        size[(size_t)r]++;
    }
    uint64_t max_clump = 0;
    for (int64_t i = 0; i < n; i++) {
        if (!em.developed[(size_t)i]) continue;
        const int64_t r = find(i);
        em.clump[(size_t)i] = id[(size_t)r];
        const int64_t s = size[(size_t)r];
        max_clump = std::max(max_clump, (uint64_t)s);
        if (s > 1) {
            const double boost = 1.15 + 0.18 * (double)(std::min<int64_t>(s, 6) - 1);
            em.fe[(size_t)i] *= boost;
        }
    }
    rep.total_clumps = (uint64_t)clumps;
    rep.max_clump_size = max_clump;
    rep.avg_clump_size = clumps > 0 ? (double)ndev / (double)clumps : 0.0;
}

}  // namespace fg

namespace fg {

void touch_graph(const Emulsion &em, double fe, double pf, std::vector<uint32_t> &ptr,
                 std::vector<uint32_t> &idx)
{
    const int64_t n = em.n;
    ptr.assign((size_t)n + 1, 0u);
    idx.clear();
    if (n == 0) return;
    std::vector<double> mdim((size_t)n);
    double maxb = 0.0;
    for (int64_t i = 0; i < n; i++) {
        const double *d = &em.dims[(size_t)(3 * i)];
        mdim[(size_t)i] = fmax(fmax(d[0], d[1]), d[2]);
        maxb = fmax(maxb, mdim[(size_t)i]);
    }
    const double cell = std::max((2.0 * maxb * fe) * pf, 1e-9);
    std::unordered_map<int64_t, std::vector<int64_t>> grid;
    auto key = [](int64_t x, int64_t y, int64_t z) {
        return (x + (1 << 20)) | ((y + (1 << 20)) << 21) | ((z + (1 << 20)) << 42);
    };
    for (int64_t i = 0; i < n; i++) {
        const double *p = &em.pos[(size_t)(3 * i)];
        grid[key((int64_t)floor(p[0] / cell), (int64_t)floor(p[1] / cell), (int64_t)floor(p[2] / cell))].push_back(i);
    }
    std::vector<uint32_t> row;
    for (int64_t i = 0; i < n; i++) {
        const double *p = &em.pos[(size_t)(3 * i)];
        const int64_t gx = (int64_t)floor(p[0] / cell), gy = (int64_t)floor(p[1] / cell),
                      gz = (int64_t)floor(p[2] / cell);
        const double ri = mdim[(size_t)i] * fe;
        row.clear();
        for (int64_t z = gz - 1; z <= gz + 1; z++)
            for (int64_t y = gy - 1; y <= gy + 1; y++)
                for (int64_t x = gx - 1; x <= gx + 1; x++) {
                    auto itc = grid.find(key(x, y, z));
                    if (itc == grid.end()) continue;
                    for (int64_t j : itc->second) {
                        if (j == i) continue;
                        const double touch = (ri + mdim[(size_t)j] * fe) * pf;
                        const double *q = &em.pos[(size_t)(3 * j)];
                        const double dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
                        if (dx * dx + dy * dy + dz * dz <= touch * touch) row.push_back((uint32_t)j);
                    }
                }
        std::sort(row.begin(), row.end());
        idx.insert(idx.end(), row.begin(), row.end());
        ptr[(size_t)i + 1] = (uint32_t)idx.size();
    }
}

}  // namespace fg
