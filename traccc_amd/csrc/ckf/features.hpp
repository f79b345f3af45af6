// The 12 branch-candidate features the scorer reads (include/traccc_amd/scorer.h):
// predicted residual pulls, chi2, layer, q/p, theta, holes, depth, chi2/ndf
// so far, the branch's standing among its seed's branches, rank among the
// branch's candidates, number of candidates passing the chi2 cut.  Computed
// from the predicted state (x, c) and the measurement, before any Kalman
// update.
//
// Standing (feature 9): the running chi2 / ndf the candidate would give its
// branch, minus the lowest such value over the rank-0 candidates of every
// branch of the same seed at this step.  The minimum is collected with a
// 64-bit atomicMin per seed keyed by the step's epoch (an older step's entry
// compares larger), so feature 9 is written as the running value first and
// the minimum subtracted by a later kernel of the step.
#pragma once

#include "device_types.hpp"
#include "kalman_device.hpp"

namespace tamd {

constexpr int kFeatures = 12;
constexpr int kFeatStanding = 9;

// running chi2 / ndf of a branch after taking a candidate
__device__ inline float branch_run(float chi2_prev, uint32_t ndf_prev, float chi2, uint32_t dims) {
    return (chi2_prev + chi2) / static_cast<float>(ndf_prev + dims);
}

__device__ inline uint64_t standing_key(uint32_t epoch, float run) {
    return (static_cast<uint64_t>(0xffffffffu - epoch) << 32) | __float_as_uint(run);
}

// Record a rank-0 candidate's running value for its seed.
__device__ inline void standing_offer(uint64_t* seed_min, uint32_t seed, uint32_t epoch, float run) {
    atomicMin(reinterpret_cast<unsigned long long*>(seed_min + seed),
              static_cast<unsigned long long>(standing_key(epoch, run)));
}

// Feature 9 from the running value and the seed's minimum of this step.
__device__ inline float standing_of(const uint64_t* seed_min, uint32_t seed, uint32_t epoch,
                                    float run) {
    const uint64_t k = seed_min[seed];
    if (static_cast<uint32_t>(k >> 32) != 0xffffffffu - epoch) {
        return 0.f;
    }
    return run - __uint_as_float(static_cast<uint32_t>(k));
}

__device__ inline void candidate_features(const ckf_params& P, const meas_data& md, const float* x,
                                          const float* c, uint32_t sf, uint32_t step,
                                          uint32_t holes, uint32_t consec, float chi2_prev,
                                          uint32_t ndf_prev, float chi2, uint32_t rank,
                                          uint32_t n_pass, float* f) {
    const proj_view v = project(x, c, md.a, md.b);
    f[0] = (md.m0 - v.x0) / sqrtf(v.p00 + md.v0);
    f[1] = md.dims == 2u ? (md.m1 - v.x1) / sqrtf(v.p11 + md.v1) : 0.f;
    f[2] = chi2;
    f[3] = static_cast<float>(P.sfs[sf].layer) / static_cast<float>(P.n_layers);
    f[4] = x[e_qop];
    f[5] = x[e_theta];
    f[6] = static_cast<float>(holes);
    f[7] = static_cast<float>(step);
    f[8] = chi2_prev / static_cast<float>(ndf_prev > 0u ? ndf_prev : 1u);
    (void)consec;
    f[kFeatStanding] = branch_run(chi2_prev, ndf_prev, chi2, md.dims);  // minus the seed's minimum later
    f[10] = static_cast<float>(rank);
    f[11] = log1pf(static_cast<float>(n_pass));
}

}  // namespace tamd
