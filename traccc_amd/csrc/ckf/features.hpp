// The 12 branch-candidate features the scorer reads (include/traccc_amd/scorer.h):
// predicted residual pulls, chi2, layer, q/p, theta, holes, depth, chi2/ndf
// so far, consecutive holes, rank among the branch's candidates, number of
// candidates passing the chi2 cut.  Computed from the predicted state (x, c)
// and the measurement, before any Kalman update.
#pragma once

#include "device_types.hpp"
#include "kalman_device.hpp"

namespace tamd {

constexpr int kFeatures = 12;

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
    f[9] = static_cast<float>(consec);
    f[10] = static_cast<float>(rank);
    f[11] = log1pf(static_cast<float>(n_pass));
}

}  // namespace tamd
