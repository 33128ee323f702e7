// Launch interfaces of the expected-model kernels (adjoint.hip): path replay of the
// photon streams for the expected dose and for its gradients, development
// probability, expected scan and its adjoint.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace fg {


// Expected-scan parameters: the BVH holds every crystal with boxes for the largest
// boosted expansion, records with semi-axes scaled by the unboosted filament expansion;
// the sens word of each record holds the crystal's transport slot.
struct ExpScanArgs {
    const Node *nodes;
    const float4 *recs;
    uint32_t n_nodes;
    uint32_t w_px, h_px, g, ss, mode;
    float pw, ph, z0, sigma;
    uint32_t key0, key1;
    uint32_t rays_per_lane, pad;   // scan shape (scan_lanes)
    const float *prob;     // development probability per transport slot
    const float *boost;    // expected clump boost of the expansion per transport slot
    float *out;            // forward: w_px * h_px
    const float *gout;     // adjoint: dL/d(out)
    double *gprob;         // adjoint: dL/d(prob) accumulators per transport slot
    double *gboost;        // adjoint: dL/d(boost) accumulators per transport slot
};

int replay_blocks(uint64_t n_photons);
// Persistent replay over resident blocks (grid = CUs x replay_persistent_blocks_per_cu());
// work must point to a zeroed 64-bit counter; the gradient pass writes grid partials.
int replay_persistent_blocks_per_cu();
hipError_t launch_replay_persistent_dose(const ReplayArgs &a, unsigned long long *work, int grid, hipStream_t s);
hipError_t launch_replay_persistent_grad(const ReplayArgs &a, unsigned long long *work, int grid, hipStream_t s);
hipError_t launch_replay_dose(const ReplayArgs &a, hipStream_t s);
hipError_t launch_replay_grad(const ReplayArgs &a, hipStream_t s);
// prob = P[Poisson(dose) >= thr] and dprob/ddose, from the float dose
hipError_t launch_dev_prob(const double *dose64, float *dose32, float *prob, float *dprob,
                           uint32_t n, uint32_t thr, hipStream_t s);
hipError_t launch_dev_prob_from(const float *dose32, float *prob, float *dprob, uint32_t n,
                                uint32_t thr, hipStream_t s);
// Expected clump boost: K_i = sum of prob over the touch neighbours (CSR in slot order),
// boost_i = B(1 + K_i) and dboost_i = B'(1 + K_i)
hipError_t launch_boost(const uint32_t *ptr, const uint32_t *idx, const float *prob, float *boost,
                        float *dboost, uint32_t n, hipStream_t s);
// gprob_j += sum over neighbours i of gboost_i * dboost_i (the touch graph is symmetric)
hipError_t launch_boost_grad(const uint32_t *ptr, const uint32_t *idx, const double *gboost,
                             const float *dboost, double *gprob, uint32_t n, hipStream_t s);
hipError_t launch_expected_scan(const ExpScanArgs &a, hipStream_t s);
hipError_t launch_expected_scan_adjoint(const ExpScanArgs &a, hipStream_t s);
// gdose = gprob * dprob (float)
hipError_t launch_dose_grad(const double *gprob, const float *dprob, float *gdose, uint32_t n,
                            hipStream_t s);
// out[j] = sum over blocks of partials[4 * b + j], j < 4, in a fixed order
hipError_t launch_reduce_partials(const double *partials, uint32_t blocks, double *out,
                                  hipStream_t s);
hipError_t launch_f64_to_f32(const double *in, float *out, uint32_t n, hipStream_t s);

}  // namespace fg
