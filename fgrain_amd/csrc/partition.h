// The in-place median partition of the reference BVH build (bvh.rs:243-256) solved per
// element, for the device builds (batched.hip) and the host tests that check it against
// the loop itself.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FG_PART_HD __host__ __device__ __forceinline__
#else
#define FG_PART_HD inline
#endif

namespace fg {

FG_PART_HD uint32_t popcount64(unsigned long long m)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint32_t)__popcll(m);
#else
    return (uint32_t)__builtin_popcountll(m);
#endif
}

// Position of the r-th (from 0) set bit of m; r < popcount(m).
FG_PART_HD uint32_t nth_set_bit(unsigned long long m, uint32_t r)
{
    uint32_t base = 0;
    for (uint32_t w = 32; w > 0; w >>= 1) {
        const uint32_t c = popcount64((m >> base) & ((1ull << w) - 1ull));
        if (r >= c) {
            r -= c;
            base += w;
        }
    }
    return base;
}

// Destination of the element at position p of a segment of `count` in the in-place
// partition of bvh.rs:243-256 (i from the start, j from the end; a small element -- its
// centroid below the midpoint -- advances i, a large one is swapped to j).  That loop
// swaps the k-th large element from the start with the k-th small one from the end
// while the first precedes the second; every large element lands at the end of the
// segment in the order the loop meets it.  Solved per element, with r the element's
// rank among its kind counted from the start and nl the number of small elements:
//   small: k = nl - r is its rank from the end; it trades places with the k-th large
//          element when p - r (large elements before it) >= k, else it stays;
//   large: when the smalls after it, nl - (p - r), number at least r, it lands at
//          count - 1 (r = 0) or one before the position of the r-th small from the end;
//          else it moves down one place.
// pos lists the smalls' positions by rank, then the larges'.  0 < nl < count.
FG_PART_HD uint32_t partition_dest(bool small, uint32_t p, uint32_t r, uint32_t nl, uint32_t count,
                                   const uint32_t *pos)
{
    if (small) {
        const uint32_t k = nl - r;
        return (p - r >= k) ? pos[nl + k - 1] : p;
    }
    if (nl - (p - r) >= r) return r == 0 ? count - 1 : pos[nl - r] - 1;
    return p - 1;
}

// The same for a segment of at most 64 elements held one per lane: ml / mr are the
// ballots of its small / large elements (bit p = position p), so the k-th element of a
// kind is the k-th set bit of its ballot.
FG_PART_HD uint32_t partition_dest64(bool small, uint32_t p, unsigned long long ml, unsigned long long mr,
                                     uint32_t count)
{
    const unsigned long long below = (1ull << p) - 1ull;
    const uint32_t nl = popcount64(ml);
    const uint32_t r = popcount64((small ? ml : mr) & below);
    if (small) {
        const uint32_t k = nl - r;
        return (p - r >= k) ? nth_set_bit(mr, k - 1) : p;
    }
    if (nl - (p - r) >= r) return r == 0 ? count - 1 : nth_set_bit(ml, nl - r) - 1;
    return p - 1;
}

}  // namespace fg
