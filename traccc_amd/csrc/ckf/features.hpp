// The 12 branch-candidate features the scorer reads (include/traccc_amd/scorer.h):
// predicted residual pulls, chi2, layer, q/p, theta, holes, depth, chi2/ndf
// so far, the branch's standing among its seed's branches, rank among the
// branch's candidates, number of candidates passing the chi2 cut.  Computed
// from the predicted state (x, c) and the measurement, before any Kalman
// update.
//
// Standing (feature 9): the branch's running chi2 / ndf (feature 8) minus the
// lowest running chi2 / ndf over the links its seed made at the previous
// step -- 0 for the seed's leading branch, positive for the ones behind it,
// 0 at step 0.  The find kernel that makes a step's links offers the lowest
// running value of each input's links to its seed's entry with a 64-bit
// atomicMin keyed by the next step's epoch (an older step's entry compares
// larger), into the half of a two-step array the next step reads; so the
// feature is final when the candidate is scored, in the same kernel.
#pragma once

#include "device_types.hpp"
#include "kalman_device.hpp"

namespace tamd {

constexpr int kFeatures = 12;
constexpr int kFeatStanding = 9;

// running chi2 / ndf of a branch (feature 8; a link's running sums)
__device__ inline float branch_run(float chi2_sum, uint32_t ndf_sum) {
    return chi2_sum / static_cast<float>(ndf_sum > 0u ? ndf_sum : 1u);
}

__device__ inline uint64_t standing_key(uint32_t epoch, float run) {
    return (static_cast<uint64_t>(0xffffffffu - epoch) << 32) | __float_as_uint(run);
}

// The seed entries step `step` reads (and step - 1 wrote).
__device__ inline const uint64_t* standing_read(const ckf_params& P, uint32_t step) {
    return P.seed_run_min + static_cast<size_t>(step & 1u) * P.n_seeds;
}
// The entries this step offers to (the next step's).
__device__ inline uint64_t* standing_write(const ckf_params& P, uint32_t step) {
    return P.seed_run_min + static_cast<size_t>((step + 1u) & 1u) * P.n_seeds;
}

// Offer the lowest running value of an input's new links for the next step
// (epoch = the next step's epoch).
__device__ inline void standing_offer(uint64_t* seed_min, uint32_t seed, uint32_t epoch, float run) {
    atomicMin(reinterpret_cast<unsigned long long*>(seed_min + seed),
              static_cast<unsigned long long>(standing_key(epoch, run)));
}

// The seed's lowest running value of the previous step's links, or -1 when
// it made none (step 0): running values are never negative.
__device__ inline float standing_min(const uint64_t* seed_min, uint32_t seed, uint32_t epoch) {
    const uint64_t k = seed_min[seed];
    if (static_cast<uint32_t>(k >> 32) != 0xffffffffu - epoch) {
        return -1.f;
    }
    return __uint_as_float(static_cast<uint32_t>(k));
}

__device__ inline void candidate_features(const ckf_params& P, const meas_data& md, const float* x,
                                          const float* c, uint32_t sf, uint32_t step,
                                          uint32_t holes, uint32_t consec, float chi2_prev,
                                          uint32_t ndf_prev, float chi2, uint32_t rank,
                                          uint32_t n_pass, float run_min, float* f) {
    const proj_view v = project(x, c, md.a, md.b);
    f[0] = (md.m0 - v.x0) / sqrtf(v.p00 + md.v0);
    f[1] = md.dims == 2u ? (md.m1 - v.x1) / sqrtf(v.p11 + md.v1) : 0.f;
    f[2] = chi2;
    f[3] = static_cast<float>(P.sfs[sf].layer) / static_cast<float>(P.n_layers);
    f[4] = x[e_qop];
    f[5] = x[e_theta];
    f[6] = static_cast<float>(holes);
    f[7] = static_cast<float>(step);
    f[8] = branch_run(chi2_prev, ndf_prev);
    (void)consec;
    f[kFeatStanding] = run_min >= 0.f ? f[8] - run_min : 0.f;
    f[10] = static_cast<float>(rank);
    f[11] = log1pf(static_cast<float>(n_pass));
}

}  // namespace tamd
