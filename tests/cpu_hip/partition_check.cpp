// Host check of the per-element form of the reference BVH partition (partition.h, used
// by the device builds) against the reference's in-place loop (bvh.rs:243-256) on random
// segments: block form (partition_dest) up to 400 elements, lane form
// (partition_dest64) up to 64.  Prints "segments N mismatches M".
#include <cstdio>
#include <random>
#include <utility>
#include <vector>

#include "partition.h"

int main()
{
    std::mt19937 g(7);
    long bad = 0, tests = 0;
    for (int t = 0; t < 200000; t++) {
        const uint32_t n = 2 + g() % (t % 2 ? 63 : 400);
        const double pr = (g() % 1001) / 1000.0;
        std::vector<int> sm(n);
        uint32_t nl = 0;
        for (auto &x : sm) {
            x = (g() % 1000) / 1000.0 < pr;
            nl += x;
        }
        if (nl == 0 || nl == n) continue;   // the build splits only when both sides are non-empty
        tests++;
        std::vector<uint32_t> a(n);
        for (uint32_t i = 0; i < n; i++) a[i] = i;
        uint32_t i = 0, j = n - 1;   // bvh.rs:243-256
        while (i <= j) {
            if (sm[a[i]]) {
                i++;
            } else {
                std::swap(a[i], a[j]);
                if (j == 0) break;
                j--;
            }
        }
        std::vector<uint32_t> pos(n), rk(n);
        uint32_t cl = 0, cb = 0;
        unsigned long long ml = 0, mr = 0;
        for (uint32_t p = 0; p < n; p++) {
            if (sm[p]) {
                rk[p] = cl;
                pos[cl++] = p;
            } else {
                rk[p] = cb;
                pos[nl + cb++] = p;
            }
            if (n <= 64) (sm[p] ? ml : mr) |= 1ull << p;
        }
        bool ok = true;
        for (uint32_t p = 0; p < n && ok; p++) {
            const uint32_t d = fg::partition_dest(sm[p], p, rk[p], nl, n, pos.data());
            ok = d < n && a[d] == p;
            if (ok && n <= 64) ok = a[fg::partition_dest64(sm[p], p, ml, mr, n)] == p;
        }
        bad += !ok;
    }
    printf("segments %ld mismatches %ld\n", tests, bad);
    return bad != 0;
}
