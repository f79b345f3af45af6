// Progressive Kalman filter (finding_config.run_pkf with one branch per seed
// or per surface): every seed follows one branch that takes the best
// measurement of each sensitive surface it reaches.
//   run_progressive_kalman_filter  core/include/traccc/finding/details/run_progressive_kalman_filter.hpp:55-258
//   progressive_kalman_filter      core/include/traccc/finding/details/progressive_kalman_filter.hpp:53-209
//   measurement_updater actor      core/include/traccc/finding/actors/measurement_kalman_updater.hpp:94-260
//   find_optimal_measurement       core/include/traccc/finding/measurement_selector.hpp:224-264
//
// There is no step structure: one lane walks one seed through the detector in
// a single persistent kernel (waves take 64-seed partitions from a ticket
// queue), writing its measurements into a fixed per-seed slot range.  A scan
// with decoupled look-back then lays the tracks out in seed order, one entry
// per seed as the reference does (an empty entry where no track survived).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "device_types.hpp"
#include "kalman_device.hpp"
#include "kernels.hpp"
#include "state_io.hpp"
#include "wave.hpp"

namespace tamd {

namespace {

constexpr int kBlock = 256;

__device__ inline bool momentum_above(const ckf_params& P, const float* x, float div) {
    const float p = P.nav.abs_q / fabsf(x[e_qop]);
    const float pT = p * sinf(x[e_theta]);
    return pT > P.nav.min_pT / div && p > P.nav.min_p / div;
}

// One seed's walk through the detector.
__device__ inline void walk_seed(const ckf_params& P, const pkf_params& Q, const nav_tables& s_nav,
                                 const nav_config<float>& leg_cfg, uint32_t s) {
    const float eps = FLT_EPSILON;
    float x[6], c[21];
    uint32_t sf = 0u;
    load_seed(P, s, x, c, sf);
    uint32_t n_states = 0u, holes = 0u, consec = 0u, ndf = 0u;
    float chi2_sum = 0.f, path = 0.f;
    bool good = false;
    if (momentum_above(P, x, 10.f)) {
        const uint32_t ev = P.seed_event[s];
        const uint32_t* rg = P.meas_ranges + static_cast<size_t>(ev) * (P.n_surfaces + 1u);
        uint32_t* states = Q.states + static_cast<size_t>(s) * Q.max_states;
        for (;;) {
            add_material(P, sf, x, c);
            // find_optimal_measurement: lowest chi2, first index on ties
            const uint32_t lo = rg[sf];
            const uint32_t up = rg[sf + 1u];
            uint32_t best = kInvalid;
            float best_chi2 = FLT_MAX;
            for (uint32_t m = lo; m < up; ++m) {
                const meas_data md = load_meas(P.meas_pos, P.meas_var, P.meas_info, m);
                const float chi2 = predicted_chi2(md, x, c);
                if (chi2 < best_chi2 && chi2 >= 0.f) {
                    best = m;
                    best_chi2 = chi2;
                    if (best_chi2 <= eps) {
                        break;
                    }
                }
            }
            int res = CKF_FIT_ERROR_OTHER;
            if (best != kInvalid && best_chi2 <= P.chi2_max) {
                const meas_data md = load_meas(P.meas_pos, P.meas_var, P.meas_info, best);
                if (best_chi2 <= eps && !(P.sfs[sf].thickness_x0 > 0.f)) {
                    // no update: take the measurement variances
                    c[sym_idx(0, 0)] = md.v0;
                    c[sym_idx(1, 1)] = md.v1;
                    res = CKF_FIT_SUCCESS;
                } else {
                    res = gain_update(md, x, c);
                }
            }
            bool exit = false;
            if (res != CKF_FIT_SUCCESS) {
                ++holes;
                ++consec;
                exit = holes > P.max_skip || consec > P.max_consec;
            } else {
                states[n_states] = best;
                ++n_states;
                consec = 0u;
                ndf += P.meas_info[best] & 0xffu;
                chi2_sum += best_chi2;
                exit = n_states >= Q.max_states;
            }
            if (exit) {
                good = true;
                break;
            }
            uint32_t n_rk = 0u;
            float leg = 0.f;
            const int status = transport_state(P, s_nav.layers, leg_cfg, x, c, sf, n_rk, leg, 0u,
                                               &s_nav, P.n_groups);
            path += leg;
            if (status == NAV_EXIT) {
                good = true;  // left the detector
                break;
            }
            if (status != NAV_FOUND || fabsf(path) > leg_cfg.path_limit) {
                break;
            }
        }
        good = good && momentum_above(P, x, 1.f) && n_states >= P.min_cands &&
               static_cast<int>(ndf) - 5 >= 0;
    }
    Q.n_states[s] = good ? n_states : 0u;
    Q.chi2[s] = chi2_sum;
    Q.ndf[s] = ndf;
    Q.holes[s] = holes;
}

__global__ __launch_bounds__(kBlock) void k_pkf(ckf_params P, pkf_params Q) {
    // the navigation layers and groups live in LDS for the whole kernel
    __shared__ nav_tables s_nav;
    stage_nav_tables(s_nav, P.layers, P.n_layers, P.groups, P.n_groups, P.nav.mask_tol_max);
    step_ctl* ctl = P.ctl;
    // legs between surfaces: no minimum step, no per-leg step budget (the
    // progressive filter's propagation has no ckf_aborter)
    nav_config<float> leg_cfg = P.nav;
    leg_cfg.min_step_length = 0.f;
    leg_cfg.max_steps = 0x7fffffffu;
    const uint32_t n = P.n_seeds;
    const uint32_t n_parts = (n + kWave - 1) / kWave;
    for (;;) {
        const uint32_t p = wave_ticket(&ctl->ticket[0].v);
        if (p >= n_parts) {
            break;
        }
        const uint32_t s = p * kWave + lane();
        if (s < n) {
            walk_seed(P, Q, s_nav, leg_cfg, s);
        }
    }
}

// Constituent offsets of the per-seed entries (seed order).
__global__ __launch_bounds__(kBlock) void k_pkf_scan(ckf_params P, pkf_params Q) {
    step_ctl* ctl = P.ctl;
    const uint32_t n = P.n_seeds;
    const uint32_t epoch = ctl->epoch;
    const uint32_t n_parts = (n + kWave - 1) / kWave;
    for (;;) {
        const uint32_t p = wave_ticket(&ctl->ticket[1].v);
        if (p >= n_parts) {
            break;
        }
        const uint32_t s = p * kWave + lane();
        const uint32_t w = s < n ? Q.n_states[s] : 0u;
        const uint32_t incl = wave_inclusive_scan(w);
        const uint32_t tot = wave_broadcast(incl, kWave - 1);
        const uint64_t b = wave_lookback(Q.lb_lo, nullptr, p, epoch, pack2(tot, 0u));
        if (s < n) {
            Q.cpos[s] = lo32(b) + incl - w;
        }
        if (p == n_parts - 1u && lane() == 0u) {
            ctl->n_tracks = n;
            ctl->n_constituents = lo32(b) + tot;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_pkf_write(ckf_params P, pkf_params Q) {
    const uint32_t n = P.n_seeds;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const uint32_t ev = P.seed_event[s];
        const uint32_t cp = Q.cpos[s];
        const uint32_t k_n = Q.n_states[s];
        if (s == 0u || P.seed_event[s - 1u] != ev) {
            P.event_const_off[ev] = cp;
        }
        const uint32_t moff = P.event_meas_off[ev];
        const uint32_t* states = Q.states + static_cast<size_t>(s) * Q.max_states;
        for (uint32_t k = 0; k < k_n; ++k) {
            P.out_const[cp + k] = states[k] - moff;
        }
        P.out_offsets[s] = cp;
        P.out_outcome[s] = 0u;
        if (k_n > 0u) {
            // track parameters are the seed's, as in traccc without a smoother
            P.out_surface[s] = P.seed_surface[s];
            copy_track_params(P.seed_vec + 6u * static_cast<size_t>(s), P.seed_cov + 36u * static_cast<size_t>(s),
                              P.out_vec + 6u * static_cast<size_t>(s), P.out_cov + 36u * static_cast<size_t>(s));
            const float ndf = static_cast<float>(static_cast<int>(Q.ndf[s]) - 5);
            P.out_ndf[s] = ndf;
            P.out_chi2[s] = Q.chi2[s];
            P.out_pval[s] = prob_chi2(Q.chi2[s], ndf);
            P.out_nholes[s] = Q.holes[s];
        } else {
            // an empty entry: default track
            P.out_surface[s] = kInvalid;
            copy_track_params(nullptr, nullptr, P.out_vec + 6u * static_cast<size_t>(s),
                              P.out_cov + 36u * static_cast<size_t>(s));
            P.out_ndf[s] = 0.f;
            P.out_chi2[s] = 0.f;
            P.out_pval[s] = 0.f;
            P.out_nholes[s] = 0u;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.out_offsets[n] = P.ctl->n_constituents;
    }
}

}  // namespace

hipError_t launch_pkf(const ckf_params& P, const pkf_params& Q, uint32_t n_cu, hipStream_t stream) {
    const uint32_t parts = (P.n_seeds + kWave - 1u) / kWave;
    const uint32_t blocks = std::max(1u, std::min((parts + 3u) / 4u, n_cu * 8u));
    hipLaunchKernelGGL(k_pkf, dim3(blocks), dim3(kBlock), 0, stream, P, Q);
    hipLaunchKernelGGL(k_pkf_scan, dim3(blocks), dim3(kBlock), 0, stream, P, Q);
    return hipGetLastError();
}

hipError_t launch_pkf_write(const ckf_params& P, const pkf_params& Q, uint32_t n_cu,
                            hipStream_t stream) {
    const uint32_t blocks =
        std::max(1u, std::min((P.n_seeds + kBlock - 1u) / kBlock, n_cu * 8u));
    hipLaunchKernelGGL(k_pkf_write, dim3(blocks), dim3(kBlock), 0, stream, P, Q);
    return hipGetLastError();
}

}  // namespace tamd
