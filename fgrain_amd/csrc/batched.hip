// Batched frame pipeline on gfx950: all tiles of a frame go through transport,
// development (threshold + clumping), the developed-crystal BVH build and the scan in
// one sequence of device kernels, captured once as a hipGraph and replayed.
//
// * transport: a persistent kernel whose waves pull 64-packet work items from one
//   device counter, tile after tile (work items are numbered tile-major, so the waves
//   in flight share one tile's nodes and records in L2);
// * development (development.rs:59-188): integer threshold; clumps as connected
//   components of the symmetric touch relation, by lock-free union-find with the root
//   at the smallest original crystal index, numbered in that order like the BFS;
// * developed-crystal BVH (transmittance.rs:82-89 -> bvh.rs:157-295): level-synchronous
//   median split.  The split sets depend only on the centroid comparison (with a
//   centroid extent above 1e-7 both sides are non-empty), so the tree, its leaf sets and
//   its boxes equal the reference's; only the order inside a leaf can differ;
// * scan: the pixels of every tile in one launch.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batched.h"
#include "partition.h"
#include "device_common.h"

namespace fg {

namespace {

using namespace dev;

constexpr int kMaxLevels = 64;
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct Seg {
    uint32_t tile, start, count, node, depth;
};

// Segments of at most kLaneSegment crystals are split by one lane (k_bvh_level_tiny),
// of at most kWaveSegment by one wave (k_bvh_level_small), larger ones by a block
// (k_bvh_level).  Each size class has its own segment list per level, so each kernel
// reads only the segments it splits.
constexpr uint32_t kLaneSegment = 16u;
constexpr uint32_t kWaveSegment = 64u;
enum : uint32_t { kSegBlock = 0, kSegWave = 1, kSegLane = 2, kSegClasses = 3 };
__host__ __device__ __forceinline__ uint32_t seg_class(uint32_t count)
{
    return count > kWaveSegment ? kSegBlock : (count > kLaneSegment ? kSegWave : kSegLane);
}

struct BNode {
    float lo[3], hi[3];
    uint32_t left, right, parent;
    uint32_t start, count, depth;
};

struct Bufs {
    // per slot
    double *pos64, *dims64, *mdim, *fe, *cen;
    uint32_t *orig, *tile_of, *devflag, *parent, *root, *csize, *root_flag, *prefix, *devpos;
    uint32_t *devflag_o;   // developed flags by crystal index (crystal_base + orig)
    uint32_t *prank, *rpos;  // partition scratch of k_bvh_level (per element)
    int32_t *clump;
    uint32_t *elem[2], *final_elem;
    float *elo, *ehi;
    // per node capacity (2 per slot)
    BNode *bn;
    uint32_t *leafcnt;
    Node *dnodes;
    ScanQNode *dqnodes;    // the scan's grid nodes, at the indices of dnodes
    PreRec *tpre;          // the transport's pre-test records, at the indices of the frame's records
    Crystal *drecs;
    float4 *dconst;        // 2 per developed-crystal record: down_constants, centre residual
    Seg *seg[2][kSegClasses];   // [level parity][size class]
    uint32_t *seg_count;        // [3 rotating levels][size class]
    uint32_t *node_count;  // per tile
    unsigned long long *work;  // transport work counter
    uint32_t *err;         // [0] depth, [1] leaf size, [2] unfinished build
    uint32_t *devlist;     // developed slots (any order) and their number
    uint32_t *dev_count;
    TileDev *tiles;
};

// ---------------------------------------------------------------- helpers -------
__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// find with path halving; parents only move towards roots, so the plain halving
// stores cannot detach a node from its tree
__device__ uint32_t uf_find(uint32_t *parent, uint32_t x)
{
    for (;;) {
        const uint32_t p = ld_relaxed(&parent[x]);
        if (p == x) return x;
        const uint32_t gp = ld_relaxed(&parent[p]);
        if (gp == p) return p;
        st_relaxed(&parent[x], gp);
        x = gp;
    }
}

// link the root with the larger original index under the smaller one
__device__ void uf_unite(uint32_t *parent, const uint32_t *orig, uint32_t a, uint32_t b)
{
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (orig[a] < orig[b]) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&parent[a], a, b) == a) return;
    }
}

__device__ __forceinline__ uint32_t tile_of_block(const TileDev *tiles, uint32_t n_tiles, uint32_t b)
{
    uint32_t lo = 0, hi = n_tiles;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tiles[mid].block_base <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------------ kernels -----
__global__ void __launch_bounds__(256) k_reset(Bufs B, uint32_t *counts, unsigned long long *stats,
                                               uint64_t n, uint32_t n_tiles)
{
    if (blockIdx.x == 0) {
        if (threadIdx.x < 3 * kSegClasses) B.seg_count[threadIdx.x] = 0u;
        if (threadIdx.x < 4) B.err[threadIdx.x] = 0u;
        if (threadIdx.x == 0) *B.dev_count = 0u;
        if (threadIdx.x == 0) *B.work = 0ull;
    }
    // counts and stats NULL: kept (development of given latent counts)
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t lim = max(2 * n, max((uint64_t)n_tiles, 8ull));
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < lim; i += stride) {
        if (i < 2 * n) B.leafcnt[i] = 0u;
        if (i < n) {
            if (counts) counts[i] = 0u;
            B.csize[i] = 0u;
            B.root_flag[i] = 0u;
        }
        if (i < n_tiles) {
            B.node_count[i] = 0u;
            B.tiles[i].n_dev = B.tiles[i].clumps = B.tiles[i].max_clump = B.tiles[i].dev_nodes = 0u;
        }
        if (i < 8 && stats) stats[i] = 0ull;
    }
}

// persistent transport over every tile's packets, built for 6 waves per SIMD (80 VGPRs)
// and for 7 (72 VGPRs, a few spills, its own LDS queue sizes): which is faster depends on
// the frame (Tri-X frame 1063 / 996 ms, T-Max frame 2921 / 3251 ms with the default queue
// sizes), so the tuning chooses
template <bool kPre, int kLQ, uint32_t kCand>
__device__ __forceinline__ void transport_frame(const TransportArgs &C, const Bufs &B, const Node *nodes,
                                                const float4 *recs, uint32_t *counts, const float *emap,
                                                uint32_t n_tiles, uint64_t total_photons)
{
    LaneCounters cnt;
    transport_persistent<true, kCounts, kPre, kLQ, kCand, true>(C, B.tiles, n_tiles, nodes, recs, counts, emap,
                                                                B.work, total_photons, cnt, nullptr, nullptr,
                                                                B.tpre);
    flush_lane_counters(cnt, C.stats);
}

// kPre: deposits pre-reduced across the wave before the global atomics (the tuning's
// transport_prereduce)
template <bool kPre>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 6))) k_transport(TransportArgs C, Bufs B, const Node *nodes,
                                                   const float4 *recs, uint32_t *counts,
                                                   const float *emap, uint32_t n_tiles,
                                                   uint64_t total_photons)
{
    transport_frame<kPre, kLeafQueue, kCandidates>(C, B, nodes, recs, counts, emap, n_tiles, total_photons);
}

template <bool kPre>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 7))) k_transport_w7(TransportArgs C, Bufs B, const Node *nodes,
                                                   const float4 *recs, uint32_t *counts,
                                                   const float *emap, uint32_t n_tiles,
                                                   uint64_t total_photons)
{
    transport_frame<kPre, kLeafQueueW7, kCandidatesW7>(C, B, nodes, recs, counts, emap, n_tiles, total_photons);
}

using TransportKernel = void (*)(TransportArgs, Bufs, const Node *, const float4 *, uint32_t *, const float *,
                                 uint32_t, uint64_t);
static TransportKernel transport_kernel_for(uint32_t waves, bool pre)
{
    if (waves == 7u) return pre ? k_transport_w7<true> : k_transport_w7<false>;
    return pre ? k_transport<true> : k_transport<false>;
}

__global__ void __launch_bounds__(256) k_threshold(Bufs B, const uint32_t *counts, uint64_t n,
                                                   DevelopParams P)
{
    __shared__ uint32_t wcnt[4], base;
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = (s < n && counts[s] >= P.threshold) ? 1u : 0u;
    if (s < n) {
        B.devflag[s] = d;
        B.devflag_o[B.tiles[B.tile_of[s]].crystal_base + B.orig[s]] = d;
        B.fe[s] = d ? P.fe : 0.0;
        B.parent[s] = (uint32_t)s;
        B.clump[s] = -1;
    }
    // append the block's developed slots to devlist (one atomic per block)
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(d != 0u);
    if (lane == 0) wcnt[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        base = tot ? atomicAdd(B.dev_count, tot) : 0u;
    }
    __syncthreads();
    if (d) {
        uint32_t off = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        for (uint32_t w = 0; w < wave; w++) off += wcnt[w];
        B.devlist[off] = (uint32_t)s;
    }
}

// Touching developed pairs within a tile, found by a box query on the tile's transport
// BVH (inclusive overlap as bvh.rs:373-379), decided in double precision; one lane per
// developed slot of devlist.  The walk queues hit leaves in an LDS column per lane and
// tests their crystals after the node steps, as the transport does; the union-find
// components do not depend on the order of the unions.
constexpr uint32_t kEdgeQueue = 8;

__global__ void __launch_bounds__(256) k_edges(Bufs B, const Node *nodes, DevelopParams P)
{
    __shared__ uint32_t lq_all[kEdgeQueue][kQueueLanes];
    uint32_t *lq = &lq_all[0][threadIdx.x];
    const uint32_t nd = *B.dev_count;
    if (blockIdx.x * blockDim.x >= nd) return;   // whole block past the list
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = q < nd;
    const uint64_t s = live ? B.devlist[q] : 0u;
    const TileDev T = B.tiles[B.tile_of[s]];
    const double px = B.pos64[3 * s], py = B.pos64[3 * s + 1], pz = B.pos64[3 * s + 2];
    const double ri = B.mdim[s] * P.fe;
    const double R = (ri + T.maxb * P.fe) * P.pf;
    const float qlo[3] = {__double2float_rd(px - R), __double2float_rd(py - R), __double2float_rd(pz - R)};
    const float qhi[3] = {__double2float_ru(px + R), __double2float_ru(py + R), __double2float_ru(pz + R)};
    // the resource is the frame's node array (wave-uniform); lanes of one wave may sit in
    // different tiles, so a lane's tile offset goes into its byte offset
    const __amdgpu_buffer_rsrc_t nodes_r = buffer_of(nodes);
    const uint32_t n_nodes = live ? T.n_nodes : 0u;
    uint32_t i = 0, nq = 0;
    for (;;) {
        const uint32_t walking = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(i < n_nodes));
        if (walking == 0u) break;
        const uint32_t stop = walking / 10u;
        uint32_t lim = n_nodes;   // 0 once the lane's queue is full
        for (uint32_t it = 0;; it++) {
            const bool stepping = i < lim;
            const uint32_t ns = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(stepping));
            if (ns == 0u || (it > 0u && ns <= stop)) break;
            if (stepping) {
                const NodeView v = load_node_buf(nodes_r, T.node_base + i);
                const bool ov = v.lo[0] <= qhi[0] && v.hi[0] >= qlo[0] && v.lo[1] <= qhi[1] &&
                                v.hi[1] >= qlo[1] && v.lo[2] <= qhi[2] && v.hi[2] >= qlo[2];
                if (ov && v.prim != kInterior) {
                    lq[kQueueLanes * nq++] = v.prim;
                    if (nq == kEdgeQueue) lim = 0u;
                }
                i = (ov && v.prim == kInterior) ? i + 1 : v.skip;
            }
        }
#pragma unroll 1
        for (uint32_t qd = 0; qd < kEdgeQueue; qd++) {
            if (qd >= nq) continue;
            const uint32_t lf = lq[kQueueLanes * qd];
            const uint32_t first = lf >> 4, cnt = lf & 15u;
            for (uint32_t k = first; k < first + cnt; k++) {
                const uint64_t j = (uint64_t)T.crystal_base + k;
                if (j <= s || !B.devflag[j]) continue;
                const double rj = B.mdim[j] * P.fe;
                const double touch = (ri + rj) * P.pf;
                const double dx = B.pos64[3 * j] - px, dy = B.pos64[3 * j + 1] - py,
                             dz = B.pos64[3 * j + 2] - pz;
                if (dx * dx + dy * dy + dz * dz <= touch * touch)
                    uf_unite(B.parent, B.orig, (uint32_t)s, (uint32_t)j);
            }
        }
        nq = 0u;
    }
}

// Each developed crystal's root goes to its own array: other threads of this kernel
// still halve paths through `parent`, so a root stored there could be overwritten by
// an ancestor.
__global__ void __launch_bounds__(256) k_roots(Bufs B, uint64_t n)
{
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || !B.devflag[s]) return;
    const uint32_t r = uf_find(B.parent, (uint32_t)s);
    B.root[s] = r;
    // one atomic per distinct root among the wave's lanes (neighbouring slots share clumps)
    for (;;) {
        const uint32_t rl = __builtin_amdgcn_readfirstlane(r);
        const unsigned long long m = __ballot(r == rl);
        if (r == rl) {
            if ((threadIdx.x & 63u) == (uint32_t)__ffsll((long long)m) - 1u)
                atomicAdd(&B.csize[rl], (uint32_t)__popcll(m));
            break;
        }
    }
    if (r == s) B.root_flag[B.tiles[B.tile_of[s]].crystal_base + B.orig[s]] = 1u;
}

// exclusive prefix of 0/1 flags over [base, base + n) by one block per tile
__device__ uint32_t block_scan_flags(const uint32_t *flags, uint32_t *out, uint32_t base, uint32_t n)
{
    __shared__ uint32_t wcnt[4];
    __shared__ uint32_t total;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t running = 0;
    for (uint32_t c = 0; c < n; c += 256) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t f = (i < n) ? flags[base + i] : 0u;
        const unsigned long long m = __ballot(f != 0u);
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t wb = 0, tot = 0;
        for (uint32_t w = 0; w < 4; w++) {
            if (w < wave) wb += wcnt[w];
            tot += wcnt[w];
        }
        if (i < n) out[base + i] = running + wb + rank;
        running += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) total = running;
    __syncthreads();
    return total;
}

__global__ void __launch_bounds__(256) k_tile_scans(Bufs B)
{
    TileDev &T = B.tiles[blockIdx.x];
    const uint32_t clumps = block_scan_flags(B.root_flag, B.prefix, T.crystal_base, T.n);
    // developed crystals numbered in crystal-index order: the developed-crystal BVH starts
    // from the subset in the order transmittance.rs:82-89 clones it
    const uint32_t ndev = block_scan_flags(B.devflag_o, B.devpos, T.crystal_base, T.n);
    if (threadIdx.x == 0) {
        T.clumps = clumps;
        T.n_dev = ndev;
    }
}

__global__ void __launch_bounds__(256) k_boost(Bufs B, uint64_t n, DevelopParams P)
{
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = s < n && B.devflag[s];
    uint32_t key = kNone, size = 0u;   // key: the tile of a clump root
    if (live) {
        const uint32_t t = B.tile_of[s];
        const uint32_t crystal_base = B.tiles[t].crystal_base;
        const uint32_t r = B.root[s];
        size = B.csize[r];
        double fe = P.fe;
        if (size > 1u) fe *= 1.15 + 0.18 * (double)(min(size, 6u) - 1u);
        B.fe[s] = fe;
        B.clump[s] = (int32_t)B.prefix[crystal_base + B.orig[r]];
        if (r == s) key = t;
        // developed-crystal box (crystal.rs:106-133 with the boosted expansion) and centroid
        const uint32_t e = crystal_base + B.devpos[crystal_base + B.orig[s]];
        B.elem[0][e] = (uint32_t)s;
        const double m = B.mdim[s] * fe;
        for (int a = 0; a < 3; a++) {
            const double lo = B.pos64[3 * s + a] - m, hi = B.pos64[3 * s + a] + m;
            B.elo[3 * s + a] = __double2float_rd(lo);
            B.ehi[3 * s + a] = __double2float_ru(hi);
            B.cen[3 * s + a] = 0.5 * (lo + hi);
        }
    }
    // largest clump per tile: every lane of the wave takes part, one atomic per tile
    // among the wave's roots
    const uint32_t lane = threadIdx.x & 63u;
    for (;;) {
        const unsigned long long pend = __ballot(key != kNone);
        if (pend == 0ull) break;
        const int leader = __ffsll((long long)pend) - 1;
        const uint32_t tl = (uint32_t)__shfl((int)key, leader, 64);
        uint32_t v = key == tl ? size : 0u;
        for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
        if ((int)lane == leader) atomicMax(&B.tiles[tl].max_clump, v);
        if (key == tl) key = kNone;
    }
}

__global__ void k_root_segments(Bufs B, uint32_t n_tiles)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    const TileDev T = B.tiles[t];
    if (T.n_dev == 0) return;
    const uint32_t c = seg_class(T.n_dev);
    const uint32_t k = atomicAdd(&B.seg_count[c], 1u);
    B.seg[0][c][k] = Seg{t, T.crystal_base, T.n_dev, 2 * T.crystal_base, 0u};
    B.node_count[t] = 1u;
    B.bn[2 * T.crystal_base].parent = kNone;
}

template <int kWaves, class V, class Op>
__device__ __forceinline__ V block_reduce(V v, V *lds, Op op)
{
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    V r = lds[0];
    for (int w = 1; w < kWaves; w++) r = op(r, lds[w]);
    return r;
}

struct MinF { __device__ float operator()(float a, float b) const { return fminf(a, b); } };
struct MaxF { __device__ float operator()(float a, float b) const { return fmaxf(a, b); } };
struct MinD { __device__ double operator()(double a, double b) const { return fmin(a, b); } };
struct MaxD { __device__ double operator()(double a, double b) const { return fmax(a, b); } };
struct SumU { __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; } };

// Destination of the element at position p of a segment of `count` in the in-place
// partition of bvh.rs:243-256 (i from the start, j from the end; a small element -- its
// centroid below the midpoint -- advances i, a large one is swapped to j).  That loop
// swaps the k-th large element from the start with the k-th small one from the end
// while the first precedes the second; every large element lands at the end of the
// segment in the order the loop meets it.  Solved per element (checked against the loop
// itself on 20 000 random segments): with r the element's rank among its kind counted
// from the start and nl the number of small elements,
//   small: k = nl - r is its rank from the end; it trades places with the k-th large
//          element when p - r (large elements before it) >= k, else it stays;
//   large: when the smalls after it, nl - (p - r), number at least r, it lands at
//          count - 1 (r = 0) or one before the position of the r-th small from the end;
//          else it moves down one place.
// pos lists the smalls' positions by rank, then the larges'.
__device__ __forceinline__ uint32_t partition_dest(bool small, uint32_t p, uint32_t r, uint32_t nl, uint32_t count,
                                                   const uint32_t *pos)
{
    if (small) {
        const uint32_t k = nl - r;
        return (p - r >= k) ? pos[nl + k - 1] : p;
    }
    if (nl - (p - r) >= r) return r == 0 ? count - 1 : pos[nl - r] - 1;
    return p - 1;
}

// Node ids of a split and its two child segments.  One lane does this per split (a block's
// thread 0, a wave's lane 0); k_bvh_level_tiny aggregates the counters over its wave.
__device__ __forceinline__ void emit_split(Bufs &B, int level, const Seg &sg, uint32_t nl, BNode &nd)
{
    const int outc = (level + 1) % 3;
    const TileDev &T = B.tiles[sg.tile];
    const uint32_t id = atomicAdd(&B.node_count[sg.tile], 2u);
    const uint32_t l = 2 * T.crystal_base + id, r = l + 1;
    nd.left = l;
    nd.right = r;
    B.bn[l].parent = sg.node;
    B.bn[r].parent = sg.node;
    if (sg.depth > 61u) atomicOr(&B.err[0], 1u);
    const Seg kids[2] = {Seg{sg.tile, sg.start, nl, l, sg.depth + 1},
                         Seg{sg.tile, sg.start + nl, sg.count - nl, r, sg.depth + 1}};
    const uint32_t c0 = seg_class(kids[0].count), c1 = seg_class(kids[1].count);
    Seg *const *out = B.seg[(level + 1) & 1];
    if (c0 == c1) {
        const uint32_t k = atomicAdd(&B.seg_count[3 * outc + c0], 2u);
        out[c0][k] = kids[0];
        out[c0][k + 1] = kids[1];
    } else {
        out[c0][atomicAdd(&B.seg_count[3 * outc + c0], 1u)] = kids[0];
        out[c1][atomicAdd(&B.seg_count[3 * outc + c1], 1u)] = kids[1];
    }
}

// One level of the median-split build (bvh.rs:166-295), one block of kLevelThreads per
// segment of more than kWaveSegment crystals.  Most such segments hold a few hundred
// crystals: 1024-lane blocks made the T-Max frame's levels 27.8 -> 265.6 ms.
constexpr uint32_t kLevelThreads = 256u, kLevelWaves = kLevelThreads / 64u;

__global__ void __launch_bounds__(kLevelThreads) k_bvh_level(Bufs B, int level)
{
    __shared__ float lf[kLevelWaves];
    __shared__ double ld[kLevelWaves];
    __shared__ uint32_t lu[kLevelWaves];
    __shared__ uint32_t wl[kLevelWaves], wr[kLevelWaves];
    const int in = level % 3, zc = (level + 2) % 3;
    const Seg *segs = B.seg[level & 1][kSegBlock];
    const uint32_t *src = B.elem[level & 1];
    uint32_t *dst = B.elem[(level + 1) & 1];
    if (blockIdx.x == 0 && threadIdx.x < kSegClasses) B.seg_count[3 * zc + threadIdx.x] = 0u;
    const uint32_t nseg = B.seg_count[3 * in + kSegBlock];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t si = blockIdx.x; si < nseg; si += gridDim.x) {
        const Seg sg = segs[si];
        float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        double cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t e = sg.start + threadIdx.x; e < sg.start + sg.count; e += kLevelThreads) {
            const uint32_t s = src[e];
            for (int a = 0; a < 3; a++) {
                blo[a] = fminf(blo[a], B.elo[3 * s + a]);
                bhi[a] = fmaxf(bhi[a], B.ehi[3 * s + a]);
                cmn[a] = fmin(cmn[a], B.cen[3 * s + a]);
                cmx[a] = fmax(cmx[a], B.cen[3 * s + a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            blo[a] = block_reduce<kLevelWaves>(blo[a], lf, MinF());
            bhi[a] = block_reduce<kLevelWaves>(bhi[a], lf, MaxF());
            cmn[a] = block_reduce<kLevelWaves>(cmn[a], ld, MinD());
            cmx[a] = block_reduce<kLevelWaves>(cmx[a], ld, MaxD());
        }
        BNode &nd = B.bn[sg.node];
        const double ex = cmx[0] - cmn[0], ey = cmx[1] - cmn[1], ez = cmx[2] - cmn[2];
        const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
        const double ext = axis == 0 ? ex : (axis == 1 ? ey : ez);
        const bool leaf = sg.count <= 4u || ext <= 1e-7;
        if (threadIdx.x == 0) {
            for (int a = 0; a < 3; a++) {
                nd.lo[a] = blo[a];
                nd.hi[a] = bhi[a];
            }
            nd.depth = sg.depth;
            nd.start = sg.start;
            nd.count = sg.count;
        }
        if (leaf) {
            for (uint32_t e = sg.start + threadIdx.x; e < sg.start + sg.count; e += kLevelThreads)
                B.final_elem[e] = src[e];
            if (threadIdx.x == 0) {
                nd.left = nd.right = kNone;
                if (sg.count > kMaxLeafCount) atomicOr(&B.err[1], 1u);
            }
            __syncthreads();
            continue;
        }
        const double mid = 0.5 * ((axis == 0 ? cmn[0] : axis == 1 ? cmn[1] : cmn[2]) +
                                  (axis == 0 ? cmx[0] : axis == 1 ? cmx[1] : cmx[2]));
        uint32_t nl = 0;
        for (uint32_t e = sg.start + threadIdx.x; e < sg.start + sg.count; e += kLevelThreads)
            nl += (B.cen[3 * src[e] + axis] < mid) ? 1u : 0u;
        nl = block_reduce<kLevelWaves>(nl, lu, SumU());
        if (nl == 0u || nl == sg.count) {
            // bvh.rs:262-266 fallback; unreachable while the centroid extent exceeds 1e-7
            for (uint32_t e = sg.start + threadIdx.x; e < sg.start + sg.count; e += kLevelThreads) dst[e] = src[e];
            nl = sg.count / 2;
        } else {
            // pass 1: each element's rank among the small (centroid < mid) or the large ones,
            // counted from the segment's start; prank lists the smalls' positions, then the
            // larges', by rank
            uint32_t base_l = 0, base_r = 0;
            for (uint32_t c = 0; c < sg.count; c += kLevelThreads) {
                const uint32_t p = c + threadIdx.x;
                const bool valid = p < sg.count;
                const uint32_t s = valid ? src[sg.start + p] : 0u;
                const bool left = valid && B.cen[3 * s + axis] < mid;
                const unsigned long long ml = __ballot(left), mr = __ballot(valid && !left);
                const unsigned long long below = (1ull << lane) - 1ull;
                if (lane == 0) {
                    wl[wave] = (uint32_t)__popcll(ml);
                    wr[wave] = (uint32_t)__popcll(mr);
                }
                __syncthreads();
                uint32_t ol = 0, orr = 0, tl = 0, tr = 0;
                for (uint32_t w = 0; w < kLevelWaves; w++) {
                    if (w < wave) {
                        ol += wl[w];
                        orr += wr[w];
                    }
                    tl += wl[w];
                    tr += wr[w];
                }
                if (valid) {
                    const uint32_t r = left ? base_l + ol + (uint32_t)__popcll(ml & below)
                                            : base_r + orr + (uint32_t)__popcll(mr & below);
                    B.prank[sg.start + (left ? r : nl + r)] = p;
                    B.rpos[sg.start + p] = r;
                }
                base_l += tl;
                base_r += tr;
                __syncthreads();
            }
            // pass 2: every element to its place in the reference's in-place partition
            for (uint32_t p = threadIdx.x; p < sg.count; p += kLevelThreads) {
                const uint32_t s = src[sg.start + p];
                const uint32_t d = partition_dest(B.cen[3 * s + axis] < mid, p, B.rpos[sg.start + p], nl, sg.count,
                                                  B.prank + sg.start);
                dst[sg.start + d] = s;
            }
        }
        if (threadIdx.x == 0) emit_split(B, level, sg, nl, nd);
        __syncthreads();
    }
}

template <class V, class Op>
__device__ __forceinline__ V wave_reduce(V v, Op op)
{
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    return v;
}

// The same level step for segments of kLaneSegment + 1 .. kWaveSegment crystals, one
// wave per segment (a lane per crystal): the box, centroid extent, split count and the
// partition are those of k_bvh_level.  A block takes four segments per
// step (block-uniform trip count) and reserves their node ids and child segments with one
// atomic per tile and per size class.
__global__ void __launch_bounds__(256) k_bvh_level_small(Bufs B, int level)
{
    __shared__ uint32_t w_tile[4], w_nl[4], w_id[4], w_k[4][2];
    __shared__ Seg w_seg[4];
    const int in = level % 3, outc = (level + 1) % 3;
    const Seg *segs = B.seg[level & 1][kSegWave];
    Seg *const *out = B.seg[(level + 1) & 1];
    const uint32_t *src = B.elem[level & 1];
    uint32_t *dst = B.elem[(level + 1) & 1];
    const uint32_t nseg = B.seg_count[3 * in + kSegWave];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t base = blockIdx.x * 4u; base < nseg; base += gridDim.x * 4u) {
        const uint32_t si = base + wave;
        bool split = false;
        if (si < nseg) {
            const Seg sg = segs[si];
            const bool valid = lane < sg.count;
            const uint32_t e = sg.start + lane;
            const uint32_t s = valid ? src[e] : 0u;
            float blo[3], bhi[3];
            double cmn[3], cmx[3], cen[3];
            for (int a = 0; a < 3; a++) {
                cen[a] = valid ? B.cen[3 * s + a] : 0.0;
                blo[a] = wave_reduce(valid ? B.elo[3 * s + a] : INFINITY, MinF());
                bhi[a] = wave_reduce(valid ? B.ehi[3 * s + a] : -INFINITY, MaxF());
                cmn[a] = wave_reduce(valid ? cen[a] : INFINITY, MinD());
                cmx[a] = wave_reduce(valid ? cen[a] : -INFINITY, MaxD());
            }
            BNode &nd = B.bn[sg.node];
            const double ex = cmx[0] - cmn[0], ey = cmx[1] - cmn[1], ez = cmx[2] - cmn[2];
            const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
            const double ext = axis == 0 ? ex : (axis == 1 ? ey : ez);
            const bool leaf = sg.count <= 4u || ext <= 1e-7;
            if (lane == 0) {
                for (int a = 0; a < 3; a++) {
                    nd.lo[a] = blo[a];
                    nd.hi[a] = bhi[a];
                }
                nd.depth = sg.depth;
                nd.start = sg.start;
                nd.count = sg.count;
            }
            if (leaf) {
                if (valid) B.final_elem[e] = s;
                if (lane == 0) {
                    nd.left = nd.right = kNone;
                    if (sg.count > kMaxLeafCount) atomicOr(&B.err[1], 1u);
                }
            } else {
                split = true;
                const double mid = 0.5 * ((axis == 0 ? cmn[0] : axis == 1 ? cmn[1] : cmn[2]) +
                                          (axis == 0 ? cmx[0] : axis == 1 ? cmx[1] : cmx[2]));
                const double c = axis == 0 ? cen[0] : (axis == 1 ? cen[1] : cen[2]);
                const bool left = valid && c < mid;
                const unsigned long long ml = __ballot(left), mr = __ballot(valid && !left);
                uint32_t nl = (uint32_t)__popcll(ml);
                if (nl == 0u || nl == sg.count) {
                    // bvh.rs:262-266 fallback; unreachable while the centroid extent exceeds 1e-7
                    if (valid) dst[e] = s;
                    nl = sg.count / 2;
                } else if (valid) {
                    dst[sg.start + partition_dest64(left, lane, ml, mr, sg.count)] = s;
                }
                if (lane == 0) {
                    w_seg[wave] = sg;
                    w_nl[wave] = nl;
                    if (sg.depth > 61u) atomicOr(&B.err[0], 1u);
                }
            }
        }
        if (lane == 0) w_tile[wave] = split ? w_seg[wave].tile : kNone;
        __syncthreads();
        if (threadIdx.x == 0) {
            // node ids per tile, child segments per class, in wave order
            uint32_t need[kSegClasses] = {0u, 0u, 0u};
            uint32_t cls[4][2];
            for (int w = 0; w < 4; w++) {
                if (w_tile[w] == kNone) continue;
                cls[w][0] = seg_class(w_nl[w]);
                cls[w][1] = seg_class(w_seg[w].count - w_nl[w]);
                need[cls[w][0]]++;
                need[cls[w][1]]++;
            }
            uint32_t kb[kSegClasses];
            for (uint32_t c = 0; c < kSegClasses; c++)
                kb[c] = need[c] ? atomicAdd(&B.seg_count[3 * outc + c], need[c]) : 0u;
            bool done[4] = {false, false, false, false};
            for (int w = 0; w < 4; w++) {
                if (w_tile[w] == kNone) continue;
                w_k[w][0] = kb[cls[w][0]]++;
                w_k[w][1] = kb[cls[w][1]]++;
                if (done[w]) continue;
                uint32_t cnt = 0;
                for (int v = w; v < 4; v++)
                    if (w_tile[v] == w_tile[w]) cnt++;
                uint32_t id = atomicAdd(&B.node_count[w_tile[w]], 2u * cnt);
                for (int v = w; v < 4; v++)
                    if (w_tile[v] == w_tile[w]) {
                        w_id[v] = id;
                        id += 2u;
                        done[v] = true;
                    }
            }
        }
        __syncthreads();
        if (split && lane == 0) {
            const Seg sg = w_seg[wave];
            const uint32_t nl = w_nl[wave];
            const uint32_t l = 2 * B.tiles[sg.tile].crystal_base + w_id[wave], r = l + 1;
            BNode &nd = B.bn[sg.node];
            nd.left = l;
            nd.right = r;
            B.bn[l].parent = sg.node;
            B.bn[r].parent = sg.node;
            out[seg_class(nl)][w_k[wave][0]] = Seg{sg.tile, sg.start, nl, l, sg.depth + 1};
            out[seg_class(sg.count - nl)][w_k[wave][1]] = Seg{sg.tile, sg.start + nl, sg.count - nl, r, sg.depth + 1};
        }
        __syncthreads();
    }
}

// The level step for segments of at most kLaneSegment crystals, one lane per segment:
// sequential box and extent (min/max are exact in any order), then the reference's
// in-place partition loop itself.  The children of such a segment are
// again at most kLaneSegment long; a wave reserves its splits' node ids (per tile) and
// child segments with one atomic each.
__global__ void __launch_bounds__(256) k_bvh_level_tiny(Bufs B, int level)
{
    const int in = level % 3, outc = (level + 1) % 3;
    const Seg *segs = B.seg[level & 1][kSegLane];
    Seg *out = B.seg[(level + 1) & 1][kSegLane];
    const uint32_t *src = B.elem[level & 1];
    uint32_t *dst = B.elem[(level + 1) & 1];
    const uint32_t nseg = B.seg_count[3 * in + kSegLane];
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t stride = gridDim.x * blockDim.x;
    // wave-uniform trip count, so that the ballots below see the whole wave
    const uint32_t first = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    for (uint32_t sw = first; sw < nseg; sw += stride) {
        const uint32_t si = sw + lane;
        bool split = false;
        Seg sg{};
        uint32_t nl = 0;
        if (si < nseg) {
            sg = segs[si];
            float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
            double cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t e = sg.start; e < sg.start + sg.count; e++) {
                const uint32_t s = src[e];
                for (int a = 0; a < 3; a++) {
                    blo[a] = fminf(blo[a], B.elo[3 * s + a]);
                    bhi[a] = fmaxf(bhi[a], B.ehi[3 * s + a]);
                    cmn[a] = fmin(cmn[a], B.cen[3 * s + a]);
                    cmx[a] = fmax(cmx[a], B.cen[3 * s + a]);
                }
            }
            BNode &nd = B.bn[sg.node];
            const double ex = cmx[0] - cmn[0], ey = cmx[1] - cmn[1], ez = cmx[2] - cmn[2];
            const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
            const double ext = axis == 0 ? ex : (axis == 1 ? ey : ez);
            const bool leaf = sg.count <= 4u || ext <= 1e-7;
            for (int a = 0; a < 3; a++) {
                nd.lo[a] = blo[a];
                nd.hi[a] = bhi[a];
            }
            nd.depth = sg.depth;
            nd.start = sg.start;
            nd.count = sg.count;
            if (leaf) {
                for (uint32_t e = sg.start; e < sg.start + sg.count; e++) B.final_elem[e] = src[e];
                nd.left = nd.right = kNone;
                if (sg.count > kMaxLeafCount) atomicOr(&B.err[1], 1u);
            } else {
                split = true;
                const double mid = 0.5 * ((axis == 0 ? cmn[0] : axis == 1 ? cmn[1] : cmn[2]) +
                                          (axis == 0 ? cmx[0] : axis == 1 ? cmx[1] : cmx[2]));
                for (uint32_t e = sg.start; e < sg.start + sg.count; e++)
                    nl += (B.cen[3 * src[e] + axis] < mid) ? 1u : 0u;
                if (nl == 0u || nl == sg.count) {
                    // bvh.rs:262-266 fallback; unreachable while the centroid extent exceeds 1e-7
                    for (uint32_t e = sg.start; e < sg.start + sg.count; e++) dst[e] = src[e];
                    nl = sg.count / 2;
                } else {
                    // the in-place partition of bvh.rs:243-256 itself, on the copy in dst
                    for (uint32_t e = sg.start; e < sg.start + sg.count; e++) dst[e] = src[e];
                    uint32_t i = sg.start, j = sg.start + sg.count - 1;
                    while (i <= j) {
                        const uint32_t s = dst[i];
                        if (B.cen[3 * s + axis] < mid) {
                            i++;
                        } else {
                            dst[i] = dst[j];
                            dst[j] = s;
                            if (j == sg.start) break;   // j would pass i (bvh.rs:251-253)
                            j--;
                        }
                    }
                }
                if (sg.depth > 61u) atomicOr(&B.err[0], 1u);
            }
        }
        // the wave's splits: two child segments each, one reservation
        const unsigned long long ms = __ballot(split);
        if (ms == 0ull) continue;
        const int leader = __ffsll((long long)ms) - 1;
        uint32_t kbase = 0;
        if ((int)lane == leader) kbase = atomicAdd(&B.seg_count[3 * outc + kSegLane], 2u * (uint32_t)__popcll(ms));
        kbase = (uint32_t)__shfl((int)kbase, leader, 64) + 2u * (uint32_t)__popcll(ms & below);
        // node ids: one reservation per tile among the wave's splits
        uint32_t key = split ? sg.tile : kNone, id = 0;
        for (;;) {
            const unsigned long long pend = __ballot(key != kNone);
            if (pend == 0ull) break;
            const int lt = __ffsll((long long)pend) - 1;
            const uint32_t tl = (uint32_t)__shfl((int)key, lt, 64);
            const unsigned long long mt = __ballot(key == tl);
            uint32_t b0 = 0;
            if ((int)lane == lt) b0 = atomicAdd(&B.node_count[tl], 2u * (uint32_t)__popcll(mt));
            b0 = (uint32_t)__shfl((int)b0, lt, 64);
            if (key == tl) {
                id = b0 + 2u * (uint32_t)__popcll(mt & below);
                key = kNone;
            }
        }
        if (split) {
            const uint32_t l = 2 * B.tiles[sg.tile].crystal_base + id, r = l + 1;
            BNode &nd = B.bn[sg.node];
            nd.left = l;
            nd.right = r;
            B.bn[l].parent = sg.node;
            B.bn[r].parent = sg.node;
            out[kbase] = Seg{sg.tile, sg.start, nl, l, sg.depth + 1};
            out[kbase + 1] = Seg{sg.tile, sg.start + nl, sg.count - nl, r, sg.depth + 1};
        }
    }
}

// slot: the segment-count slot the level after the last launched one reads
// ((last + 1) % 3, which is kMaxLevels % 3 when every level was launched)
__global__ void k_build_check(Bufs B, uint32_t slot)
{
    const uint32_t *left = B.seg_count + 3u * slot;
    if (threadIdx.x == 0 && (left[0] | left[1] | left[2]) != 0u) B.err[2] = 1u;
}

__device__ __forceinline__ bool node_valid(const Bufs &B, uint64_t g, uint32_t &t)
{
    t = B.tile_of[g >> 1];
    const TileDev &T = B.tiles[t];
    return g - 2ull * T.crystal_base < B.node_count[t];
}

// Leaves under every node, bottom-up from the leaves: each child adds its count plus
// kArrive to its parent's (zeroed) counter; the first child to arrive stops, the second
// reads the sibling's count from the atomic's return value and carries the sum upward.
// Both additions of kArrive wrap to zero, so every counter ends at its leaf count.  One
// uncontended atomic per child, instead of one per leaf on every ancestor.
constexpr uint32_t kArrive = 0x80000000u;

__global__ void __launch_bounds__(256) k_leafcount(Bufs B, uint64_t cap)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t t;
    if (g >= cap || !node_valid(B, g, t)) return;
    if (B.bn[g].left != kNone) return;
    B.leafcnt[g] = 1u;
    uint32_t c = 1u;
    for (uint32_t x = B.bn[g].parent; x != kNone; x = B.bn[x].parent) {
        const uint32_t old = atomicAdd(&B.leafcnt[x], c + kArrive);
        if (!(old & kArrive)) break;
        c += old - kArrive;
    }
}

// threaded position and skip of every node: a right child sits right after its
// parent, a left child after its right sibling's subtree (2 * leaves - 1 nodes); a
// leaf ends its right-child ancestor's subtree (or the tree), so its skip is slot + 1
__global__ void __launch_bounds__(256) k_layout(Bufs B, uint64_t cap)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t t;
    if (g >= cap || !node_valid(B, g, t)) return;
    TileDev &T = B.tiles[t];
    const uint32_t nn = B.node_count[t];
    uint32_t off = 0, rel = 0, ysize = 0;
    bool found = false;
    uint32_t x = (uint32_t)g;
    for (;;) {
        const uint32_t p = B.bn[x].parent;
        if (p == kNone) break;
        const uint32_t r = B.bn[p].right;
        if (r == x) {
            if (!found) {
                found = true;
                rel = off;
                ysize = 2 * B.leafcnt[x] - 1;
            }
            off += 1;
        } else {
            off += 1 + (2 * B.leafcnt[r] - 1);
        }
        x = p;
    }
    const BNode &nd = B.bn[g];
    Node o;
    for (int a = 0; a < 3; a++) {
        o.lo[a] = nd.lo[a];
        o.hi[a] = nd.hi[a];
    }
    o.skip = found ? (off - rel) + ysize : nn;
    o.prim = (nd.left == kNone) ? ((nd.start - T.crystal_base) << 4) | min(nd.count, kMaxLeafCount) : kInterior;
    B.dnodes[2ull * T.crystal_base + off] = o;
    if (g == 2ull * T.crystal_base) T.dev_nodes = nn;  // root: one writer per tile
}

// the scan's 16-byte grid node of every node of the developed-crystal trees
__global__ void __launch_bounds__(256) k_scan_qnodes(Bufs B, uint64_t cap)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t t;
    if (g >= cap || !node_valid(B, g, t)) return;
    float bx, by, sc;
    scan_qframe(B.dnodes[2ull * B.tiles[t].crystal_base], bx, by, sc);
    B.dqnodes[g] = scan_qnode(B.dnodes[g], bx, by, sc);
}

// the transport's pre-test record of every crystal of the frame (PreRec)
__global__ void __launch_bounds__(256) k_pre_records(Bufs B, const Crystal *recs, uint64_t n)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Crystal r = recs[k];
    B.tpre[k] = pre_record(r.c, r.h, r.morph);
}

__global__ void __launch_bounds__(256) k_dev_records(Bufs B, const Crystal *recs, uint64_t n)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const TileDev &T = B.tiles[B.tile_of[e]];
    if (e - T.crystal_base >= T.n_dev) return;
    const uint32_t s = B.final_elem[e];
    Crystal r = recs[s];
    const double fe = B.fe[s];
    for (int a = 0; a < 3; a++) r.h[a] = (float)(B.dims64[3 * (size_t)s + a] * fe);
    float k[4];
    down_constants(r.q, r.h, r.morph, k);
    r.sens = k[3];   // the scan reads the bounding-sphere radius with the centre
    B.drecs[e] = r;
    B.dconst[2 * e] = make_float4(k[0], k[1], k[2], k[3]);
    const double *p = &B.pos64[3 * (size_t)s];
    B.dconst[2 * e + 1] = make_float4((float)(p[0] - (double)r.c[0]), (float)(p[1] - (double)r.c[1]),
                                      (float)(p[2] - (double)r.c[2]), 0.0f);
}

__global__ void __launch_bounds__(256) k_scan(Bufs B, ScanCommon S, float *out, uint32_t n_tiles)
{
    const uint32_t t = tile_of_block(B.tiles, n_tiles, blockIdx.x);
    const TileDev &T = B.tiles[t];
    const uint32_t L = scan_lanes(S.g * S.g * S.ss * S.ss, S.rays_per_lane);
    const uint32_t pix = (blockIdx.x - T.block_base) * (256u / L) + threadIdx.x / L;
    const bool valid = pix < T.tw * T.th;
    const uint32_t px = valid ? pix % T.tw : 0u, py = valid ? pix / T.tw : 0u;
    const Node *nodes = B.dnodes + 2ull * T.crystal_base;
    const ScanQNode *qnodes = B.dqnodes + 2ull * T.crystal_base;
    const float4 *recs = reinterpret_cast<const float4 *>(B.drecs + T.crystal_base);
    const float Tm = scan_pixel(valid, pix, px, py, threadIdx.x % L, L, S.g, S.ss, T.pw, T.ph, S.z0,
                                S.sigma, T.skey0, T.skey1, nodes, qnodes, B.node_count[t], recs,
                                B.dconst + 2ull * T.crystal_base);
    if (valid && threadIdx.x % L == 0u)
        out[T.out_base + (size_t)py * T.out_stride + px] = scan_value(Tm, S.mode);
}

// ------------------------------------------------- transport BVHs on the device
// BVH::build (bvh.rs:157-195) of every crystal of each tile: element boxes by
// HalideCrystal::aabb of an undeveloped crystal (crystal.rs:106-133, expansion 1),
// rounded outward to f32, f64 centroids; slots are in crystal order within a tile, so
// the build starts from the index order 0..n like the reference.
__global__ void __launch_bounds__(256) k_tb_init(Bufs B, uint64_t n, uint32_t n_tiles)
{
    if (blockIdx.x == 0) {
        if (threadIdx.x < 3 * kSegClasses) B.seg_count[threadIdx.x] = 0u;
        if (threadIdx.x < 4) B.err[threadIdx.x] = 0u;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < max(2 * n, (uint64_t)n_tiles);
         i += stride) {
        if (i < 2 * n) B.leafcnt[i] = 0u;
        if (i < n_tiles) B.node_count[i] = 0u;
        if (i >= n) continue;
        uint32_t lo = 0, hi = n_tiles;   // last tile with crystal_base <= i
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (B.tiles[mid].crystal_base <= i) lo = mid;
            else hi = mid;
        }
        const uint32_t s = (uint32_t)i;
        B.tile_of[s] = lo;
        B.orig[s] = s - B.tiles[lo].crystal_base;
        const double *d = &B.dims64[3 * (size_t)s];
        const double m = fmax(fmax(d[0], d[1]), d[2]);
        B.mdim[s] = m;
        for (int a = 0; a < 3; a++) {
            const double l = B.pos64[3 * (size_t)s + a] - m, h = B.pos64[3 * (size_t)s + a] + m;
            B.elo[3 * (size_t)s + a] = __double2float_rd(l);
            B.ehi[3 * (size_t)s + a] = __double2float_ru(h);
            B.cen[3 * (size_t)s + a] = 0.5 * (l + h);
        }
        B.elem[0][s] = s;
    }
}

// The batched input's slot arrays from device-resident crystals: slot sl of tile t holds
// the tile's crystal leaf[sl] (an index within the tile), whose attributes sit at
// crystal_base + leaf[sl] of pos / dims (crystal order).
__global__ void __launch_bounds__(256) k_gather_slots(Bufs B, const double *pos, const double *dims,
                                                      const uint32_t *leaf, uint64_t n, uint32_t n_tiles)
{
    const uint64_t sl = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= n) return;
    uint32_t lo = 0, hi = n_tiles;   // last tile with crystal_base <= sl
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (B.tiles[mid].crystal_base <= sl) lo = mid;
        else hi = mid;
    }
    const uint32_t local = leaf[sl];
    const size_t cs = (size_t)B.tiles[lo].crystal_base + local;
    double m = 0.0;
    for (int a = 0; a < 3; a++) {
        B.pos64[3 * sl + a] = pos[3 * cs + a];
        B.dims64[3 * sl + a] = dims[3 * cs + a];
        m = fmax(m, dims[3 * cs + a]);
    }
    B.mdim[sl] = m;
    B.orig[sl] = local;
    B.tile_of[sl] = lo;
}

// transport records in leaf order (HalideCrystal fields as f32, make_records in host.cpp)
__global__ void __launch_bounds__(256) k_tb_records(Bufs B, const double *quat, const double *sens, uint32_t morph,
                                                    Crystal *recs, uint32_t *leaf_local, uint64_t n)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t s = B.final_elem[e];
    Crystal r;
    for (int a = 0; a < 3; a++) {
        r.c[a] = (float)B.pos64[3 * (size_t)s + a];
        r.h[a] = (float)B.dims64[3 * (size_t)s + a];
    }
    for (int a = 0; a < 4; a++) r.q[a] = (float)quat[4 * (size_t)s + a];
    r.sens = (float)sens[s];
    r.morph = morph;
    recs[e] = r;
    leaf_local[e] = B.orig[s];   // the crystal's index within its tile
}

}  // namespace

// --------------------------------------------------------------- host side -------
struct BatchedFrame {
    Bufs B{};
    std::vector<void *> allocs;
    BatchedInput in;
    uint32_t n_tiles = 0;
    uint64_t total_items = 0;
    uint64_t total_photons = 0;
    uint32_t total_blocks = 0;
    int transport_grid = 0;
    // one captured graph per stage (transport; development + developed-crystal BVH; scan)
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // stage brackets
    const char *graph_step = "";   // the graph call that failed last (batched_graph_failure)
};

template <class T>
static hipError_t alloc(BatchedFrame *b, T *&p, size_t count)
{
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    b->allocs.push_back(v);
    p = reinterpret_cast<T *>(v);
    return hipSuccess;
}

#define TRY(x)                       \
    do {                             \
        hipError_t e_ = (x);         \
        if (e_ != hipSuccess) return e_; \
    } while (0)

hipError_t batched_create(const BatchedInput &in, hipStream_t s, BatchedFrame **out)
{
    BatchedFrame *b = new BatchedFrame();
    *out = b;
    b->in = in;
    b->n_tiles = (uint32_t)in.tiles.size();
    const size_t n = (size_t)in.n_crystals;
    Bufs &B = b->B;
    TRY(alloc(b, B.pos64, 3 * n));
    TRY(alloc(b, B.dims64, 3 * n));
    TRY(alloc(b, B.mdim, n));
    TRY(alloc(b, B.fe, n));
    TRY(alloc(b, B.cen, 3 * n));
    TRY(alloc(b, B.orig, n));
    TRY(alloc(b, B.tile_of, n));
    TRY(alloc(b, B.devflag, n));
    TRY(alloc(b, B.parent, n));
    TRY(alloc(b, B.root, n));
    TRY(alloc(b, B.csize, n));
    TRY(alloc(b, B.root_flag, n));
    TRY(alloc(b, B.prefix, n));
    TRY(alloc(b, B.devpos, n));
    TRY(alloc(b, B.devflag_o, n));
    TRY(alloc(b, B.prank, n));
    TRY(alloc(b, B.rpos, n));
    TRY(alloc(b, B.clump, n));
    TRY(alloc(b, B.elem[0], n));
    TRY(alloc(b, B.elem[1], n));
    TRY(alloc(b, B.final_elem, n));
    TRY(alloc(b, B.elo, 3 * n));
    TRY(alloc(b, B.ehi, 3 * n));
    TRY(alloc(b, B.bn, 2 * n));
    TRY(alloc(b, B.leafcnt, 2 * n));
    TRY(alloc(b, B.dnodes, 2 * n));
    TRY(alloc(b, B.dqnodes, 2 * n));
    TRY(alloc(b, B.drecs, n));
    TRY(alloc(b, B.dconst, 2 * n));
    // a level's segments partition its developed crystals: at most n / 65 have more
    // than kWaveSegment crystals, n / 17 more than kLaneSegment
    for (int p = 0; p < 2; p++) {
        TRY(alloc(b, B.seg[p][kSegBlock], n / (kWaveSegment + 1) + 1));
        TRY(alloc(b, B.seg[p][kSegWave], n / (kLaneSegment + 1) + 1));
        TRY(alloc(b, B.seg[p][kSegLane], n));
    }
    TRY(alloc(b, B.seg_count, 3 * kSegClasses));
    TRY(alloc(b, B.node_count, b->n_tiles));
    TRY(alloc(b, B.work, 1));  // 64-bit packet counter
    TRY(alloc(b, B.err, 4));
    TRY(alloc(b, B.devlist, n));
    TRY(alloc(b, B.dev_count, 1));
    TRY(alloc(b, B.tiles, b->n_tiles));
    uint64_t n_trecs = 1;
    for (const TileDev &t : in.tiles) n_trecs = std::max<uint64_t>(n_trecs, (uint64_t)t.crystal_base + t.n);
    TRY(alloc(b, B.tpre, n_trecs));
    // host-side per-slot arrays
    std::vector<double> mdim(n);
    std::vector<uint32_t> tile_of(n);
    std::vector<TileDev> tiles = in.tiles;
    uint64_t items = 0;
    uint32_t blocks = 0;
    for (uint32_t t = 0; t < b->n_tiles; t++) {
        TileDev &T = tiles[t];
        T.item_base = items;
        items += (T.n_photons + 63) / 64;
        T.photon_base = b->total_photons;
        b->total_photons += T.n_photons;
        const uint32_t ppb = 256u / scan_lanes(in.scan.g * in.scan.g * in.scan.ss * in.scan.ss,
                                               in.scan.rays_per_lane);
        T.blocks_x = 0;
        T.blocks = (uint32_t)(((uint64_t)T.tw * T.th + ppb - 1) / ppb);
        T.block_base = blocks;
        blocks += T.blocks;
        if (in.dev_pos) continue;   // T.maxb as given (gathered on the device below)
        double mx = 0.0;
        for (uint32_t k = 0; k < T.n; k++) {
            const size_t sl = (size_t)T.crystal_base + k;
            const double *d = &in.dims64[3 * sl];
            mdim[sl] = fmax(fmax(d[0], d[1]), d[2]);
            mx = fmax(mx, mdim[sl]);
            tile_of[sl] = t;
        }
        T.maxb = mx;
    }
    b->total_items = items;
    b->total_blocks = blocks;
    b->in.tiles = tiles;
    TRY(hipMemcpyAsync(B.tiles, tiles.data(), tiles.size() * sizeof(TileDev), hipMemcpyHostToDevice, s));
    if (in.recs) {
        // the transport's pre-test records, once: the frame's records do not change after it is made
        (void)hipGetLastError();
        hipLaunchKernelGGL(k_pre_records, dim3((uint32_t)((n_trecs + 255) / 256)), dim3(256), 0, s, B, in.recs,
                           n_trecs);
        TRY(hipGetLastError());
    }
    if (in.dev_pos) {
        (void)hipGetLastError();
        hipLaunchKernelGGL(k_gather_slots, dim3(std::max<uint32_t>(1u, (uint32_t)((n + 255) / 256))), dim3(256), 0, s,
                           B, in.dev_pos, in.dev_dims, in.dev_leaf, (uint64_t)n, b->n_tiles);
        TRY(hipGetLastError());
    } else {
        TRY(hipMemcpyAsync(B.pos64, in.pos64.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        TRY(hipMemcpyAsync(B.dims64, in.dims64.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        TRY(hipMemcpyAsync(B.mdim, mdim.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
        TRY(hipMemcpyAsync(B.orig, in.orig.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        TRY(hipMemcpyAsync(B.tile_of, tile_of.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    }
    TRY(hipStreamSynchronize(s));
    int dev = 0, cus = 0, per_cu = 0;
    TRY(hipGetDevice(&dev));
    TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (in.transport_waves != 6u && in.transport_waves != 7u) return hipErrorInvalidValue;
    const void *tk = reinterpret_cast<const void *>(transport_kernel_for(in.transport_waves, in.transport.prereduce != 0u));
    TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tk, 256, 0));
    b->transport_grid = std::max(1, cus * std::max(per_cu, 1));
    for (auto &e : b->ev) TRY(hipEventCreate(&e));
    return hipSuccess;
}

// A failed launch or event record of the sequence names itself in b->graph_step.
#define TRY_STEP(b, what, x)           \
    do {                               \
        hipError_t e_ = (x);           \
        if (e_ != hipSuccess) {        \
            (b)->graph_step = (what);  \
            return e_;                 \
        }                              \
    } while (0)
#define TRY_LAUNCH(b, name) TRY_STEP(b, name, hipGetLastError())

// The level kernels of a BVH build, level 0 up.  A capture records all kMaxLevels levels
// (the graph is replayed for any tree); direct launches read the segment counts back every
// 8 levels and stop once a level leaves no segments (trees end after 20-30 levels and each
// empty level is three launches).  check_slot is the slot k_build_check reads.
static hipError_t launch_levels(BatchedFrame *b, Bufs &B, hipStream_t s, uint32_t &check_slot)
{
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    TRY(hipStreamIsCapturing(s, &cap));
    const bool can_stop = cap == hipStreamCaptureStatusNone;
    check_slot = kMaxLevels % 3;
    for (int level = 0; level < kMaxLevels; level++) {
        hipLaunchKernelGGL(k_bvh_level, dim3(2048), dim3(kLevelThreads), 0, s, B, level);
        if (b) TRY_LAUNCH(b, "k_bvh_level");
        hipLaunchKernelGGL(k_bvh_level_small, dim3(2048), dim3(256), 0, s, B, level);
        if (b) TRY_LAUNCH(b, "k_bvh_level_small");
        hipLaunchKernelGGL(k_bvh_level_tiny, dim3(2048), dim3(256), 0, s, B, level);
        if (b) TRY_LAUNCH(b, "k_bvh_level_tiny");
        if (can_stop && level % 8 == 7 && level + 1 < kMaxLevels) {
            uint32_t sc[3 * kSegClasses];
            TRY(hipMemcpyAsync(sc, B.seg_count, sizeof sc, hipMemcpyDeviceToHost, s));
            TRY(hipStreamSynchronize(s));
            const uint32_t next = (uint32_t)(level + 1) % 3u;
            if ((sc[3 * next] | sc[3 * next + 1] | sc[3 * next + 2]) == 0u) {
                check_slot = next;
                break;
            }
        }
    }
    return hipSuccess;
}

// Stage 0: reset and transport; 1: development and the developed-crystal BVH; 2: scan;
// kStageResetKeepCounts: the reset of stage 0 without the latent counts and counters.
constexpr int kStageResetKeepCounts = 3;
static hipError_t launch_stage(BatchedFrame *b, int stage, hipStream_t s)
{
    // every launch is checked through hipGetLastError: clear what an earlier, unrelated
    // call of this thread may have left there, or a stale error would read as a failed
    // launch (and a failed capture)
    (void)hipGetLastError();
    const BatchedInput &in = b->in;
    Bufs &B = b->B;
    const uint64_t n = (uint64_t)in.n_crystals;
    // at least one block per launch: a tile may hold no crystal
    const uint32_t nb = std::max<uint32_t>(1u, (uint32_t)((n + 255) / 256));
    const uint32_t nb2 = std::max<uint32_t>(1u, (uint32_t)((2 * n + 255) / 256));
    if (stage == 0 || stage == kStageResetKeepCounts) {
        hipLaunchKernelGGL(k_reset, dim3(std::min<uint32_t>(nb2, 8192)), dim3(256), 0, s, B,
                           stage == 0 ? in.counts : nullptr, stage == 0 ? in.stats : nullptr, n, b->n_tiles);
        TRY_LAUNCH(b, "k_reset");
        if (stage == kStageResetKeepCounts) return hipSuccess;
        TransportArgs C = in.transport;
        C.stats = in.stats;
        C.estride = in.frame_w;
        hipLaunchKernelGGL(transport_kernel_for(in.transport_waves, C.prereduce != 0u), dim3(b->transport_grid),
                           dim3(256), 0, s, C, B, in.nodes, reinterpret_cast<const float4 *>(in.recs),
                           in.counts, in.emap, b->n_tiles, b->total_photons);
        TRY_LAUNCH(b, "k_transport");
        return hipSuccess;
    }
    if (stage == 1) {
        hipLaunchKernelGGL(k_threshold, dim3(nb), dim3(256), 0, s, B, in.counts, n, in.develop);
        TRY_LAUNCH(b, "k_threshold");
        hipLaunchKernelGGL(k_edges, dim3(nb), dim3(256), 0, s, B, in.nodes, in.develop);
        TRY_LAUNCH(b, "k_edges");
        hipLaunchKernelGGL(k_roots, dim3(nb), dim3(256), 0, s, B, n);
        TRY_LAUNCH(b, "k_roots");
        hipLaunchKernelGGL(k_tile_scans, dim3(b->n_tiles), dim3(256), 0, s, B);
        TRY_LAUNCH(b, "k_tile_scans");
        hipLaunchKernelGGL(k_boost, dim3(nb), dim3(256), 0, s, B, n, in.develop);
        TRY_LAUNCH(b, "k_boost");
        hipLaunchKernelGGL(k_root_segments, dim3((b->n_tiles + 255) / 256), dim3(256), 0, s, B, b->n_tiles);
        TRY_LAUNCH(b, "k_root_segments");
    uint32_t check_slot = kMaxLevels % 3;
    TRY(launch_levels(b, B, s, check_slot));
    hipLaunchKernelGGL(k_build_check, dim3(1), dim3(64), 0, s, B, check_slot);
    TRY_LAUNCH(b, "k_build_check");
    hipLaunchKernelGGL(k_leafcount, dim3(nb2), dim3(256), 0, s, B, 2 * n);
    TRY_LAUNCH(b, "k_leafcount");
    hipLaunchKernelGGL(k_layout, dim3(nb2), dim3(256), 0, s, B, 2 * n);
    TRY_LAUNCH(b, "k_layout");
    hipLaunchKernelGGL(k_scan_qnodes, dim3(nb2), dim3(256), 0, s, B, 2 * n);
    TRY_LAUNCH(b, "k_scan_qnodes");
    hipLaunchKernelGGL(k_dev_records, dim3(nb), dim3(256), 0, s, B, in.recs, n);
    TRY_LAUNCH(b, "k_dev_records");
    return hipSuccess;
    }
    hipLaunchKernelGGL(k_scan, dim3(b->total_blocks), dim3(256), 0, s, B, in.scan, in.dneg, b->n_tiles);
    TRY_LAUNCH(b, "k_scan");
    return hipSuccess;
}

// Stage k of the render as a graph, captured on first use; a failed capture keeps
// nothing (the caller falls back to direct launches).
static hipError_t stage_graph(BatchedFrame *b, int k, hipStream_t s)
{
    if (!b->exec[k]) {
        if (b->graph[k]) {
            (void)hipGraphDestroy(b->graph[k]);
            b->graph[k] = nullptr;
        }
        b->graph_step = "hipStreamBeginCapture";
        TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t e = launch_stage(b, k, s);   // names its failing launch in graph_step
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e == hipSuccess && e2 != hipSuccess) {
            e = e2;
            b->graph_step = "hipStreamEndCapture";
        }
        if (e == hipSuccess) {
            e = hipGraphInstantiate(&b->exec[k], g, nullptr, nullptr, 0);
            if (e != hipSuccess) b->graph_step = "hipGraphInstantiate";
        }
        if (e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            b->exec[k] = nullptr;
            return e;
        }
        b->graph[k] = g;
    }
    b->graph_step = "hipGraphLaunch";
    return hipGraphLaunch(b->exec[k], s);
}

// The three stages bracketed by the timing events, which are recorded on the stream
// between the stages rather than inside a graph: event-record nodes of captured graphs
// are not available with every HIP runtime a process may load (the one bundled with
// PyTorch rejects them).
hipError_t batched_render(BatchedFrame *b, hipStream_t s, bool use_graph)
{
    TRY(hipEventRecord(b->ev[0], s));
    for (int k = 0; k < 3; k++) {
        TRY(use_graph ? stage_graph(b, k, s) : launch_stage(b, k, s));
        TRY(hipEventRecord(b->ev[k + 1], s));
    }
    return hipSuccess;
}

hipError_t batched_develop(BatchedFrame *b, hipStream_t s)
{
    TRY(launch_stage(b, kStageResetKeepCounts, s));
    TRY(hipEventRecord(b->ev[0], s));
    TRY(hipEventRecord(b->ev[1], s));
    TRY(launch_stage(b, 1, s));
    TRY(hipEventRecord(b->ev[2], s));
    return hipEventRecord(b->ev[3], s);
}

const char *batched_graph_failure(const BatchedFrame *b) { return b ? b->graph_step : ""; }

hipError_t batched_read_tiles(BatchedFrame *b, hipStream_t s, std::vector<TileDev> &tiles, uint32_t *err)
{
    tiles.resize(b->n_tiles);
    TRY(hipMemcpyAsync(tiles.data(), b->B.tiles, tiles.size() * sizeof(TileDev), hipMemcpyDeviceToHost, s));
    TRY(hipMemcpyAsync(err, b->B.err, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

hipError_t batched_read_develop(BatchedFrame *b, hipStream_t s, std::vector<uint8_t> &dev,
                                std::vector<double> &fe, std::vector<int32_t> &clump)
{
    const size_t n = (size_t)b->in.n_crystals;
    std::vector<uint32_t> d(n);
    fe.resize(n);
    clump.resize(n);
    TRY(hipMemcpyAsync(d.data(), b->B.devflag, n * 4, hipMemcpyDeviceToHost, s));
    TRY(hipMemcpyAsync(fe.data(), b->B.fe, n * 8, hipMemcpyDeviceToHost, s));
    TRY(hipMemcpyAsync(clump.data(), b->B.clump, n * 4, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    dev.resize(n);
    for (size_t i = 0; i < n; i++) dev[i] = (uint8_t)d[i];
    return hipSuccess;
}

hipError_t batched_kernel_ms(BatchedFrame *b, float *transport, float *develop, float *scan)
{
    TRY(hipEventSynchronize(b->ev[3]));
    TRY(hipEventElapsedTime(transport, b->ev[0], b->ev[1]));
    TRY(hipEventElapsedTime(develop, b->ev[1], b->ev[2]));
    return hipEventElapsedTime(scan, b->ev[2], b->ev[3]);
}

hipError_t build_transport_bvhs(const TransportBuildIn &in, Node *nodes, Crystal *recs, uint32_t *leaf_slot,
                                std::vector<uint32_t> &node_count, uint32_t err[4], hipStream_t s)
{
    const size_t n = (size_t)in.n;
    const uint32_t n_tiles = (uint32_t)in.base.size();
    node_count.assign(n_tiles, 0u);
    for (int k = 0; k < 4; k++) err[k] = 0u;
    if (n == 0 || n_tiles == 0) return hipSuccess;
    BatchedFrame tmp;   // holds the build's scratch allocations (freed below)
    Bufs &B = tmp.B;
    B.pos64 = const_cast<double *>(in.pos);
    B.dims64 = const_cast<double *>(in.dims);
    B.dnodes = nodes;
    hipError_t e = hipSuccess;
    auto A = [&](auto *&p, size_t count) {
        if (e == hipSuccess) e = alloc(&tmp, p, count);
    };
    A(B.mdim, n);
    A(B.cen, 3 * n);
    A(B.orig, n);
    A(B.tile_of, n);
    A(B.elo, 3 * n);
    A(B.ehi, 3 * n);
    A(B.elem[0], n);
    A(B.elem[1], n);
    A(B.final_elem, n);
    A(B.bn, 2 * n);
    A(B.leafcnt, 2 * n);
    A(B.prank, n);
    A(B.rpos, n);
    for (int p = 0; p < 2; p++) {
        A(B.seg[p][kSegBlock], n / (kWaveSegment + 1) + 1);
        A(B.seg[p][kSegWave], n / (kLaneSegment + 1) + 1);
        A(B.seg[p][kSegLane], n);
    }
    A(B.seg_count, 3 * kSegClasses);
    A(B.node_count, n_tiles);
    A(B.err, 4);
    A(B.tiles, n_tiles);
    std::vector<TileDev> tiles(n_tiles);
    for (uint32_t t = 0; t < n_tiles; t++) {
        tiles[t] = TileDev{};
        tiles[t].crystal_base = in.base[t];
        tiles[t].n = tiles[t].n_dev = in.count[t];
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(B.tiles, tiles.data(), n_tiles * sizeof(TileDev), hipMemcpyHostToDevice, s);
    const uint32_t nb = (uint32_t)((n + 255) / 256), nb2 = (uint32_t)((2 * n + 255) / 256);
    (void)hipGetLastError();
    uint32_t check_slot = kMaxLevels % 3;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_tb_init, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(nb2, 8192))), dim3(256), 0, s,
                           B, (uint64_t)n, n_tiles);
        hipLaunchKernelGGL(k_root_segments, dim3((n_tiles + 255) / 256), dim3(256), 0, s, B, n_tiles);
        e = launch_levels(nullptr, B, s, check_slot);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_build_check, dim3(1), dim3(64), 0, s, B, check_slot);
        hipLaunchKernelGGL(k_leafcount, dim3(nb2), dim3(256), 0, s, B, (uint64_t)(2 * n));
        hipLaunchKernelGGL(k_layout, dim3(nb2), dim3(256), 0, s, B, (uint64_t)(2 * n));
        hipLaunchKernelGGL(k_tb_records, dim3(nb), dim3(256), 0, s, B, in.quat, in.sens, in.morph, recs, leaf_slot,
                           (uint64_t)n);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(node_count.data(), B.node_count, n_tiles * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(err, B.err, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e == hipSuccess) e = e2;
    for (void *p : tmp.allocs) (void)hipFree(p);
    tmp.allocs.clear();
    return e;
}

void batched_destroy(BatchedFrame *b)
{
    if (!b) return;
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (int k = 0; k < 3; k++) {
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    }
    for (void *p : b->allocs) (void)hipFree(p);
    delete b;
}

}  // namespace fg
