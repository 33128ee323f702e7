// Host check of the scan's 16-bit grid nodes (ScanQNode, fg_layout.h) against the float
// box test they stand in for (a ray along -z enters a node when its origin lies strictly
// inside the node's x-y box): scan_q_may must hold for every point strictly inside the
// box and scan_q_surely for none outside it.  Built with -ffp-contract=off from the same
// header as the kernels.  Trees from 0.01 to 3000 units across at offsets up to 2000;
// node boxes anywhere inside the root (often sharing its edges, often tiny); points
// within a few float ulps of the box edges, inside and around the box, and far outside
// the tree.
// Usage: scan_grid_check SAMPLES SEED THREADS; prints
// "samples N inside I surely S band B violations V".
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "fg_layout.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u() { return (double)(next() >> 11) * 0x1p-53; }
    double range(double a, double b) { return a + (b - a) * u(); }
    double logu(double a, double b) { return std::exp(range(std::log(a), std::log(b))); }
    int pick(int n) { return (int)(next() % (uint64_t)n); }
};

float step_ulps(float v, int k)
{
    for (; k > 0; k--) v = std::nextafterf(v, INFINITY);
    for (; k < 0; k++) v = std::nextafterf(v, -INFINITY);
    return v;
}

// a coordinate of the test point on one axis, given the node's interval [lo, hi] and the
// tree's [rlo, rhi]
float coord(Rng &r, float lo, float hi, float rlo, float rhi)
{
    const double ext = (double)rhi - (double)rlo;
    switch (r.pick(5)) {
    case 0: return step_ulps(lo, r.pick(9) - 4);
    case 1: return step_ulps(hi, r.pick(9) - 4);
    case 2: {
        const double w = ((double)hi - (double)lo) + 1e-3 * ext;
        return (float)r.range((double)lo - w, (double)hi + w);
    }
    case 3: return (float)r.range((double)lo, (double)hi);
    default: return (float)r.range((double)rlo - 0.5 * ext, (double)rhi + 0.5 * ext);
    }
}

struct Counts {
    uint64_t n = 0, inside = 0, surely = 0, band = 0, bad = 0;
};

void run(uint64_t samples, uint64_t seed, Counts &c, std::atomic<int> &printed)
{
    Rng r{seed};
    for (uint64_t k = 0; k < samples; k++) {
        fg::Node root{}, nd{};
        for (int a = 0; a < 2; a++) {
            root.lo[a] = (float)r.range(-2000.0, 2000.0);
            root.hi[a] = (float)((double)root.lo[a] + r.logu(1e-2, 3000.0));
            if (!(root.hi[a] > root.lo[a])) root.hi[a] = std::nextafterf(root.lo[a], INFINITY);
            float a0 = (float)r.range(root.lo[a], root.hi[a]), a1 = (float)r.range(root.lo[a], root.hi[a]);
            if (a1 < a0) std::swap(a0, a1);
            if (r.u() < 0.3) a1 = std::fmin(root.hi[a], (float)((double)a0 + r.logu(1e-7, 1.0) * ((double)root.hi[a] - a0)));
            if (r.u() < 0.2) a0 = root.lo[a];
            if (r.u() < 0.2) a1 = root.hi[a];
            nd.lo[a] = a0;
            nd.hi[a] = a1;
        }
        nd.lo[2] = root.lo[2] = 0.0f;
        nd.hi[2] = root.hi[2] = 1.0f;
        nd.skip = 1;
        nd.prim = fg::kInterior;
        float bx, by, s;
        fg::scan_qframe(root, bx, by, s);
        const fg::ScanQNode q = fg::scan_qnode(nd, bx, by, s);
        const float ox = coord(r, nd.lo[0], nd.hi[0], root.lo[0], root.hi[0]);
        const float oy = coord(r, nd.lo[1], nd.hi[1], root.lo[1], root.hi[1]);
        const bool inside = nd.lo[0] < ox && ox < nd.hi[0] && nd.lo[1] < oy && oy < nd.hi[1];
        const fg::ScanQRay qr = fg::scan_qray(ox, oy, bx, by, s);
        const bool may = fg::scan_q_may(q.lo, q.hi, qr), surely = fg::scan_q_surely(q.lo, q.hi, qr);
        c.n++;
        c.inside += inside;
        c.surely += surely;
        c.band += inside && !surely;
        if ((inside && !may) || (surely && !inside)) {
            c.bad++;
            if (printed.fetch_add(1) < 5)
                std::printf("violation: box [%.9g %.9g] x [%.9g %.9g] root [%.9g %.9g] x [%.9g %.9g] point (%.9g %.9g)"
                            " inside %d may %d surely %d\n",
                            nd.lo[0], nd.hi[0], nd.lo[1], nd.hi[1], root.lo[0], root.hi[0], root.lo[1], root.hi[1],
                            ox, oy, inside, may, surely);
        }
    }
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s SAMPLES SEED THREADS\n", argv[0]);
        return 2;
    }
    const uint64_t samples = std::strtoull(argv[1], nullptr, 10), seed = std::strtoull(argv[2], nullptr, 10);
    const int threads = std::max(1, std::atoi(argv[3]));
    std::vector<Counts> c(threads);
    std::vector<std::thread> pool;
    std::atomic<int> printed{0};
    for (int t = 0; t < threads; t++)
        pool.emplace_back(run, samples / threads, seed * 1000003ull + t, std::ref(c[t]), std::ref(printed));
    for (auto &th : pool) th.join();
    Counts all;
    for (const Counts &x : c) {
        all.n += x.n;
        all.inside += x.inside;
        all.surely += x.surely;
        all.band += x.band;
        all.bad += x.bad;
    }
    std::printf("samples %llu inside %llu surely %llu band %llu violations %llu\n", (unsigned long long)all.n,
                (unsigned long long)all.inside, (unsigned long long)all.surely, (unsigned long long)all.band,
                (unsigned long long)all.bad);
    return all.bad ? 1 : 0;
}
