// Host check of the transport's absorption pre-test (crystal_may_absorb, fg_layout.h),
// and of its packed record (crystal_may_absorb_pre: it must keep every segment the float
// pre-test keeps),
// against the float crystal test itself (crystal_hit and the event conditions of the
// transport, transport.rs:109-131): a pre-test that rejects a segment on which the float
// test records an event is a violation.  Built with -ffp-contract=off from the same header
// as the kernels, so crystal_hit returns the kernels' bits.  Samples favour the cases
// where float error is largest: origins far from the crystal (up to 3e4 bounding radii),
// rays aimed within 1e-8 .. 1 radii of the bounding sphere and of the crystal's surface,
// segment ends next to the crystal, flat ellipsoids (aspect up to 40) and rays with a zero
// world component.
// Usage: absorb_pretest_check SAMPLES SEED THREADS; prints
// "samples N events E rejected R violations V".
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
    void unit(double v[3])
    {
        for (;;) {
            double n = 0;
            for (int i = 0; i < 3; i++) {
                v[i] = range(-1, 1);
                n += v[i] * v[i];
            }
            if (n > 1e-6 && n <= 1) {
                n = std::sqrt(n);
                for (int i = 0; i < 3; i++) v[i] /= n;
                return;
            }
        }
    }
};

// q v q^-1 for the rotation whose inverse crystal_hit applies (crystal frame -> world)
void rotate(const double q[4], const double v[3], double out[3])
{
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double tx = 2 * (y * v[2] - z * v[1]), ty = 2 * (z * v[0] - x * v[2]), tz = 2 * (x * v[1] - y * v[0]);
    out[0] = v[0] + w * tx + (y * tz - z * ty);
    out[1] = v[1] + w * ty + (z * tx - x * tz);
    out[2] = v[2] + w * tz + (x * ty - y * tx);
}

struct Totals {
    std::atomic<long> n{0}, events{0}, rejected{0}, bad{0};
};

void run(uint64_t seed, long samples, Totals &T)
{
    Rng g{seed};
    long n = 0, ev = 0, rej = 0, bad = 0;
    for (long s = 0; s < samples; s++) {
        const uint32_t morph = (uint32_t)(g.next() % 3);
        const double r = g.logu(0.02, 3.0);
        double h[3] = {r, r, r};
        if (morph == 1) h[2] = std::max(r / g.logu(1.0, 40.0), 0.005);
        if (g.u() < 0.2)
            for (double &x : h) x = r * g.logu(0.1, 1.0);
        double qd[4], qn = 0;
        for (double &x : qd) {
            x = g.range(-1, 1);
            qn += x * x;
        }
        qn = std::sqrt(qn);
        for (double &x : qd) x /= qn;
        const double hm = std::max(std::max(h[0], h[1]), h[2]);
        const double Rb = morph == 0 ? std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]) : hm;
        double c[3];
        for (double &x : c) x = g.range(0, 3000);
        // a target point near the bounding sphere or near the crystal's surface
        double tgt[3], w[3];
        g.unit(w);
        const int mode = (int)(g.next() % 4);
        if (mode == 0) {
            const double f = g.u() < 0.5 ? 1 + g.logu(1e-8, 1.0) : 1 - g.logu(1e-8, 1.0);
            for (int i = 0; i < 3; i++) tgt[i] = c[i] + w[i] * Rb * f;
        } else if (mode == 1) {
            // a point of the crystal's surface (crystal frame), pushed out or in by a little
            double l[3];
            if (morph == 0) {
                const int ax = (int)(g.next() % 3);
                for (int i = 0; i < 3; i++) l[i] = g.range(-h[i], h[i]);
                l[ax] = (g.u() < 0.5 ? -1 : 1) * h[ax];
            } else {
                for (int i = 0; i < 3; i++) l[i] = w[i] * h[i];
                const double m = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                double nrm = 0;
                for (int i = 0; i < 3; i++) nrm += (l[i] / h[i]) * (l[i] / h[i]);
                for (int i = 0; i < 3; i++) l[i] /= std::sqrt(nrm) / m;
            }
            const double f = 1 + (g.u() < 0.5 ? 1 : -1) * g.logu(1e-9, 1e-2);
            for (double &x : l) x *= f;
            double wl[3];
            rotate(qd, l, wl);
            for (int i = 0; i < 3; i++) tgt[i] = c[i] + wl[i];
        } else {
            for (int i = 0; i < 3; i++) tgt[i] = c[i] + w[i] * Rb * g.range(0, 3);
        }
        // origin at distance |p| from the target, direction towards it (or at random)
        double u[3];
        g.unit(u);
        const double dist = Rb * g.logu(1e-3, 3e4);
        double o[3], d[3];
        for (int i = 0; i < 3; i++) o[i] = tgt[i] - u[i] * dist;
        if (mode == 3) g.unit(d);
        else
            for (int i = 0; i < 3; i++) d[i] = u[i];
        if (g.u() < 0.05) d[g.next() % 3] = 0.0;
        float cf[3], hf[3], qf[4];
        for (int i = 0; i < 3; i++) {
            cf[i] = (float)c[i];
            hf[i] = (float)h[i];
        }
        for (int i = 0; i < 4; i++) qf[i] = (float)qd[i];
        const float ox = (float)o[0], oy = (float)o[1], oz = (float)o[2];
        float dx = (float)d[0], dy = (float)d[1], dz = (float)d[2];
        const float dn = sqrtf(dx * dx + dy * dy + dz * dz);
        if (!(dn > 0.0f)) continue;
        dx /= dn;
        dy /= dn;
        dz /= dn;
        // the segment ends next to the crystal's extent along the ray, or anywhere
        const double tp = (c[0] - ox) * dx + (c[1] - oy) * dy + (c[2] - oz) * dz;
        double st;
        const int sm = (int)(g.next() % 3);
        if (sm == 0) st = tp + Rb * (g.u() < 0.5 ? 1 : -1) * (1 + (g.u() < 0.5 ? 1 : -1) * g.logu(1e-9, 1.0));
        else if (sm == 1) st = tp + (g.u() < 0.5 ? 1 : -1) * Rb * g.logu(1e-9, 2.0);
        else st = Rb * g.logu(1e-4, 1e5);
        const float step = (float)std::max(st, 1e-7);
        float t0 = 0, t1 = 0;
        const bool hit = fg::crystal_hit(cf, qf, hf, morph, ox, oy, oz, dx, dy, dz, t0, t1);
        const bool event = hit && t0 < step && fmaxf(fminf(t1, step) - t0, 0.0f) > 0.0f;
        const bool may = fg::crystal_may_absorb(cf, hf, morph, ox, oy, oz, dx, dy, dz, step);
        // the frame transport's packed pre-test record must keep whatever the float one keeps
        const bool may_pre = fg::crystal_may_absorb_pre(fg::pre_record(cf, hf, morph), ox, oy, oz, dx, dy, dz, step);
        n++;
        ev += event;
        rej += !may;
        if ((event && !may) || (may && !may_pre)) {
            if (bad < 5)
                std::printf("violation%s: morph %u h %.9g %.9g %.9g c %.9g %.9g %.9g o %.9g %.9g %.9g d %.9g %.9g %.9g "
                            "step %.9g t0 %.9g t1 %.9g\n",
                            may ? " (packed record)" : "", morph, hf[0], hf[1], hf[2], cf[0], cf[1], cf[2], ox, oy, oz,
                            dx, dy, dz, step, t0, t1);
            bad++;
        }
    }
    T.n += n;
    T.events += ev;
    T.rejected += rej;
    T.bad += bad;
}

}  // namespace

int main(int argc, char **argv)
{
    const long samples = argc > 1 ? std::atol(argv[1]) : 2000000;
    const uint64_t seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
    const int threads = argc > 3 ? std::atoi(argv[3]) : 1;
    Totals T;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back(run, seed * 1000003ull + (uint64_t)t, samples / threads, std::ref(T));
    for (auto &th : pool) th.join();
    std::printf("samples %ld events %ld rejected %ld violations %ld\n", T.n.load(), T.events.load(),
                T.rejected.load(), T.bad.load());
    return T.bad.load() == 0 ? 0 : 1;
}
