// Internal host-side structures of the MI355X engine (not part of the C ABI).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fgrain_hip.h"
#include "fg_layout.h"

namespace fg {

// BVH in the reference's own layout (bvh.rs:129-151): pre-order with the left child at
// node + 1, f64 boxes, leaf ranges into `idx`.
struct RefBvh {
    std::vector<double> mn, mx;  // 3 per node
    std::vector<uint32_t> left, right, offset, count;
    std::vector<uint32_t> idx;   // crystal index per leaf slot
    int64_t nodes() const { return (int64_t)left.size(); }
};

struct DeviceEmulsion;  // device copies (engine.cpp)

struct Emulsion {
    fg_emulsion_config cfg{};
    int64_t n = 0;
    std::vector<double> pos, dims, quat, sens;  // 3, 3, 4, 1 per crystal
    std::vector<int32_t> morph;
    // latent image and development state, crystal order
    std::vector<uint32_t> counts;
    std::vector<uint8_t> developed;
    std::vector<double> fe;
    std::vector<int64_t> clump;
    // BVH over all crystals with undeveloped boxes (built by synthesize, emulsion.rs:213)
    RefBvh bvh;
    std::vector<Node> nodes;      // threaded device layout of bvh
    std::vector<Crystal> recs;    // records in leaf order (bvh.idx order)
    DeviceEmulsion *dev = nullptr;
    bool dev_state_dirty = true;  // counts changed on the host since the last upload
    double reach = 0.0;           // (sqrt(3) - 1) * largest semi-axis of a cubic crystal
    // a tile of a frame built on the device: pos / dims / quat / sens stay in the frame's
    // device arrays (empty here), rmax bounds the largest semi-axis
    bool on_device = false;
    double rmax = 0.0;
};

// Philox sequential stream with the layout documented in fg_layout.h.
struct SeqRng {
    uint32_t key[2];
    uint64_t block = 0;
    uint32_t buf[4];
    int pos = 4;
    explicit SeqRng(uint64_t seed) { key[0] = (uint32_t)seed; key[1] = (uint32_t)(seed >> 32); }
    uint32_t u32();
    uint64_t u64() { const uint64_t lo = u32(); const uint64_t hi = u32(); return lo | (hi << 32); }
    double uniform() { return (double)(u64() >> 11) * 0x1p-53; }
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)u64() * n) >> 64); }
    double normal(double mean, double sd);
};

// host.cpp
void synthesize(const fg_emulsion_config &cfg, uint64_t seed, Emulsion &em);
// attributes of em.pos's points (em.n set), each crystal from its own Philox stream
// (device synthesis, whose points arrive in cell order)
void attributes_per_point(const fg_emulsion_config &cfg, uint64_t seed, Emulsion &em);
void build_ref_bvh(int64_t n, const double *pos, const double *dims, const uint8_t *developed,
                   const double *fe, const uint32_t *subset, RefBvh &out);
// Threaded layout of a reference BVH; false (with msg) if a leaf holds more than 15
// crystals or the tree is deeper than the reference's 64-entry stack can hold.
bool flatten_bvh(const RefBvh &b, std::vector<Node> &out, std::string &msg);
void make_records(const Emulsion &em, const RefBvh &b, const uint8_t *developed,
                  const double *fe, std::vector<Crystal> &out);
bool finalize_emulsion(Emulsion &em, std::string &msg);
float round_down_f(double x);
float round_up_f(double x);
// Touch graph of development.rs:132-155 with unboosted expansion fe: j is a neighbour of
// i when |p_i - p_j| <= (m_i + m_j) * fe * pf (m = largest semi-axis); CSR in crystal
// order, rows sorted.
void touch_graph(const Emulsion &em, double fe, double pf, std::vector<uint32_t> &ptr,
                 std::vector<uint32_t> &idx);
// threshold + clumping on the host from em.counts (union-find over a uniform grid)
void develop_host(Emulsion &em, const fg_development_config &cfg, fg_development_report &rep);

}  // namespace fg
