// Kalman track fit with the two-filter smoother on CDNA: the C ABI of
// include/traccc_amd/fitting.h and its kernels.
//
// What it replaces: traccc's device kalman_fitting_algorithm
// (device/common/src/fitting/kalman_fitting_algorithm.cpp: fit_prelude, then
// fit_forward / fit_backward kernels per iteration) around
// kalman_fitter::filter / smooth
// (core/include/traccc/fitting/kalman_filter/kalman_fitter.hpp:177-420),
// kalman_actor (kalman_actor.hpp) and two_filters_smoother
// (two_filters_smoother.hpp:38-294).
//
// One lane fits one track candidate.  k_fit_forward walks the navigation from
// the seed surface: material first, then the gain-matrix update on the
// candidate's measurement of the surface; any other sensitive surface is a
// hole (kept in the track's surface sequence when it carries material).
// k_fit_backward starts from the last filtered state with the covariance
// inflated and walks the sequence back with direct plane-to-plane transport;
// on every state it combines the backward prediction with the forward
// filtered state, then runs the backward filter, then the material.
//
// The smoother's combination is the reference's weighted mean
//   C_s = (C_p^-1 + C_f^-1)^-1,  x_s = C_s (C_f^-1 x_f + C_p^-1 x_p)
// written in its equivalent gain form: with S = C_f + C_p = L L^T
// (Cholesky), Y = L^-1 C_f and z = L^-1 (x_p - x_f),
//   C_s = C_f - Y^T Y,  x_s = x_f + Y^T z,
// which needs one 6x6 factorisation instead of three inverses and keeps FP32
// away from the cancellation of the information form.  Before combining,
// x_p's phi is moved by 2 pi to within pi of x_f's (the reference combines
// raw angles, wrong at phi ~ +-pi).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "../ckf/handles.hpp"
#include "../ckf/kalman_device.hpp"
#include "../ckf/state_io.hpp"
#include "../common/errors.hpp"
#include "traccc_amd/fitting.h"

namespace tamd {
namespace {

constexpr uint32_t kFitBlock = 128;
constexpr uint32_t kSeqState = 0x80000000u;  // sequence entry is a track state

// kalman_fitter_status values (core/include/traccc/fitting/status_codes.hpp:14-29)
enum : uint32_t {
    kSuccess = 0u,
    kQopZero = 1u,
    kThetaPole = 2u,
    kInversion = 3u,
    kSmootherChi2Negative = 4u,
    kSmootherChi2NotFinite = 5u,
    kSmootherInvalidCov = 6u,
    kSmootherSkipped = 7u,
    kUpdaterInvalidCov = 8u,
    kUpdaterSkipped = 9u,
    kSequenceOverflow = 10u,
    kPropagationFailure = 11u,
    kIncomplete = 0xffu  // navigation left the detector before the last state
};

// Filtered / smoothed record of one state: x[6], packed C[21], two chi2 words,
// surface, flags -- 128 B, one record per lane and state.
constexpr int kRec = 32;
constexpr int kRecChi2a = 27;  // filtered chi2 | smoothed chi2
constexpr int kRecChi2b = 28;  // -            | backward chi2

struct fit_params {
    const ckf_surface* sfs;
    const ckf_layer* layers;
    uint32_t n_layers;
    const layer_group* groups;
    uint32_t n_groups;
    field_view<float> field;
    nav_config<float> nav;
    float inflation;
    uint32_t seq_factor, seq_min;
    // events
    const float2* meas_pos;
    const float2* meas_var;
    const uint32_t* meas_info;
    const uint32_t* meas_sf;
    const uint32_t* ev_meas_off;   // [n_events]
    const uint32_t* ev_track_off;  // [n_events + 1]
    uint32_t n_events;
    // candidates
    uint32_t n_tracks;
    const uint32_t* cand_surface;
    const float* cand_vec;   // [n][6]
    const float* cand_cov;   // [n][36]
    const uint32_t* offsets; // [n + 1]
    const uint32_t* cons;    // event-relative measurement index
    // work + output
    uint32_t* seq;           // [offsets[t] * factor + t * min .. )
    uint32_t* fw_info;       // [n] status | holes << 8 | n_seq << 16
    float* flt;              // [n_states][kRec]
    float* smt;              // [n_states][kRec]
    uint8_t* flags;          // [n_states]
    uint32_t* state_sf;      // [n_states] surface index of each state
    uint16_t* outcome;       // [n]
    uint32_t* out_surface;   // [n]
    float* out_vec;          // [n][6]
    float* out_cov;          // [n][21]
    float* out_ndf;
    float* out_chi2;
    float* out_pval;
    uint32_t* out_nholes;
};

__device__ inline uint32_t track_event(const fit_params& F, uint32_t t) {
    uint32_t lo = 0u, hi = F.n_events;  // largest e with ev_track_off[e] <= t
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) / 2u;
        if (F.ev_track_off[mid] <= t) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return lo;
}

__device__ inline void inflate(float* c, float f) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = i; j < 6; ++j) {
            c[sym_idx(i, j)] = (i == j) ? c[sym_idx(i, i)] * f : 0.f;
        }
    }
}

__device__ inline void store_rec(float* r, const float* x, const float* c) {
#pragma unroll
    for (int k = 0; k < 6; ++k) r[k] = x[k];
#pragma unroll
    for (int k = 0; k < 21; ++k) r[6 + k] = c[k];
}

__device__ inline void load_rec(const float* r, float* x, float* c) {
#pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = r[k];
#pragma unroll
    for (int k = 0; k < 21; ++k) c[k] = r[6 + k];
}

// chi2 of residual m - H x against R = V - H C H^T (a filtered or smoothed
// state: the measurement minus the state's own projected covariance); 1-D
// measurements use R00 only (masked_inverse).
__device__ inline float residual_chi2(const meas_data& m, const float* x, const float* c) {
    const proj_view v = project(x, c, m.a, m.b);
    const float r0 = m.m0 - v.x0;
    const float R00 = m.v0 - v.p00;
    if (m.dims == 1u) {
        return r0 * (1.f / R00) * r0;
    }
    const float r1 = m.m1 - v.x1;
    const float R01 = -v.p01;
    const float R11 = m.v1 - v.p11;
    const float det = R00 * R11 - R01 * R01;
    const float i00 = R11 / det, i11 = R00 / det, i01 = -R01 / det;
    return r0 * (i00 * r0 + i01 * r1) + r1 * (i01 * r0 + i11 * r1);
}

// Smoother on one state: (xp, cp) backward prediction, (xf, cf) forward
// filtered; on success (xs, cs) smoothed, the two chi2 values, and (xp, cp)
// replaced by the backward filtered state.
__device__ inline uint32_t smooth_state(const meas_data& m, float* xp, float* cp, const float* xf,
                                        const float* cf, float* xs, float* cs, float& chi2_s,
                                        float& chi2_b) {
    // S = C_f + C_p = L L^T
    float L[21];
#pragma unroll
    for (int k = 0; k < 21; ++k) L[k] = cf[k] + cp[k];
    // in-place Cholesky on the packed upper triangle: L[sym_idx(j, i)] holds
    // element (i, j) of the lower factor, i >= j
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        float d = L[sym_idx(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= L[sym_idx(k, j)] * L[sym_idx(k, j)];
        if (!(d > 0.f) || !isfinite(d)) {
            return kInversion;
        }
        const float ljj = sqrtf(d);
        const float inv = 1.f / ljj;
        L[sym_idx(j, j)] = ljj;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            float s = L[sym_idx(j, i)];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[sym_idx(k, i)] * L[sym_idx(k, j)];
            L[sym_idx(j, i)] = s * inv;
        }
    }
    // Y = L^-1 C_f (forward substitution per column), z = L^-1 (x_p - x_f)
    float Y[6][6];
    float z[6];
    float d[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = xp[k] - xf[k];
    {
        const float pi = 3.14159265358979323846f, twopi = 6.28318530717958647692f;
        if (d[e_phi] > pi) d[e_phi] -= twopi;
        if (d[e_phi] < -pi) d[e_phi] += twopi;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float inv = 1.f / L[sym_idx(i, i)];
        float zi = d[i];
#pragma unroll
        for (int k = 0; k < i; ++k) zi -= L[sym_idx(k, i)] * z[k];
        z[i] = zi * inv;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            float y = cf[sym_idx(i, c)];
#pragma unroll
            for (int k = 0; k < i; ++k) y -= L[sym_idx(k, i)] * Y[k][c];
            Y[i][c] = y * inv;
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        float acc = xf[i];
#pragma unroll
        for (int k = 0; k < 6; ++k) acc += Y[k][i] * z[k];
        xs[i] = acc;
#pragma unroll
        for (int j = i; j < 6; ++j) {
            float s = cf[sym_idx(i, j)];
#pragma unroll
            for (int k = 0; k < 6; ++k) s -= Y[k][i] * Y[k][j];
            cs[sym_idx(i, j)] = s;
        }
    }
    // regularize_covariance (min variance -0.01)
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        if (cs[sym_idx(i, i)] < -0.01f) {
            return kSmootherInvalidCov;
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) cs[sym_idx(i, i)] = fabsf(cs[sym_idx(i, i)]);
    if (!isfinite(xs[e_theta]) || !isfinite(xs[e_phi])) {
        return kInversion;
    }
    if (fabsf(xs[e_qop]) == 0.f) {
        return kQopZero;
    }
    if (!normalize_angles(xs[e_phi], xs[e_theta])) {
        return kThetaPole;
    }
    chi2_s = residual_chi2(m, xs, cs);
    if (chi2_s < -10.f) {
        return kSmootherChi2Negative;
    }
    if (!isfinite(chi2_s)) {
        return kSmootherChi2NotFinite;
    }
    // backward filter: the gain update of the backward prediction
    float xb[6], cb[21];
#pragma unroll
    for (int k = 0; k < 6; ++k) xb[k] = xp[k];
#pragma unroll
    for (int k = 0; k < 21; ++k) cb[k] = cp[k];
    const int st = gain_update(m, xb, cb);
    if (st != CKF_FIT_SUCCESS) {
        return st == CKF_FIT_ERROR_UPDATER_INVALID_COVARIANCE ? kSmootherInvalidCov
                                                              : static_cast<uint32_t>(st);
    }
    chi2_b = residual_chi2(m, xb, cb);
    if (chi2_b < 0.f) {
        return kSmootherChi2Negative;
    }
    if (!isfinite(chi2_b)) {
        return kSmootherChi2NotFinite;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) xp[k] = xb[k];
#pragma unroll
    for (int k = 0; k < 21; ++k) cp[k] = cb[k];
    return kSuccess;
}

__global__ __launch_bounds__(kFitBlock) void k_fit_forward(fit_params F, uint32_t iteration) {
    __shared__ nav_tables s_nav;
    stage_nav_tables(s_nav, F.layers, F.n_layers, F.groups, F.n_groups, F.nav.mask_tol_max);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= F.n_tracks) {
        return;
    }
    if (iteration > 0u && F.outcome[t] != CKF_OUTCOME_SUCCESS) {
        return;  // a failed iteration ends the fit of this track
    }
    const uint32_t s0 = F.offsets[t];
    const uint32_t n = F.offsets[t + 1u] - s0;
    const uint32_t moff = F.ev_meas_off[track_event(F, t)];
    const geo_view G{F.sfs, F.n_layers, F.field};
    float x[6], c[21];
    uint32_t sf = 0u;
    if (iteration == 0u) {
#pragma unroll
        for (int k = 0; k < 6; ++k) x[k] = F.cand_vec[6u * static_cast<size_t>(t) + k];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
#pragma unroll
            for (int j = i; j < 6; ++j) c[sym_idx(i, j)] = F.cand_cov[36u * static_cast<size_t>(t) + 6 * i + j];
        }
        sf = F.cand_surface[t];
    } else {
        // the first state's filtered parameters start the next iteration
        load_rec(F.flt + static_cast<size_t>(s0) * kRec, x, c);
        sf = F.meas_sf[moff + F.cons[s0]];
    }
    inflate(c, F.inflation);
    for (uint32_t k = 0; k < n; ++k) {
        F.flags[s0 + k] = 0u;
        F.state_sf[s0 + k] = F.meas_sf[moff + F.cons[s0 + k]];
    }
    uint32_t* seq = F.seq + static_cast<size_t>(s0) * F.seq_factor + static_cast<size_t>(t) * F.seq_min;
    const uint32_t cap = max(n * F.seq_factor, F.seq_min);
    uint32_t k = 0u, holes = 0u, n_seq = 0u, status = n > 0u ? kSuccess : kIncomplete;
    uint32_t next_sf = n > 0u ? F.meas_sf[moff + F.cons[s0]] : 0xffffffffu;
    while (k < n) {
        add_material_sf(F.sfs[sf], F.nav.mass, F.nav.abs_q, x, c);
        if (sf == next_sf) {
            const uint32_t mi = moff + F.cons[s0 + k];
            const meas_data md = load_meas(F.meas_pos, F.meas_var, F.meas_info, mi);
            const int st = gain_update(md, x, c);
            if (st != CKF_FIT_SUCCESS) {
                status = static_cast<uint32_t>(st);
                break;
            }
            float* r = F.flt + static_cast<size_t>(s0 + k) * kRec;
            store_rec(r, x, c);
            r[kRecChi2a] = predicted_chi2(md, x, c);
            if (n_seq == cap) {
                status = kSequenceOverflow;
                break;
            }
            seq[n_seq++] = kSeqState | sf;
            if (++k == n) {
                break;
            }
            next_sf = F.meas_sf[moff + F.cons[s0 + k]];
        } else {
            bool later = false;
            for (uint32_t j = k + 1u; j < n; ++j) {
                later = later || (F.meas_sf[moff + F.cons[s0 + j]] == sf);
            }
            if (later) {
                status = kUpdaterSkipped;
                break;
            }
            ++holes;
            if (F.sfs[sf].thickness_x0 > 0.f) {
                if (n_seq == cap) {
                    status = kSequenceOverflow;
                    break;
                }
                seq[n_seq++] = sf;
            }
        }
        uint32_t n_rk = 0u;
        float leg = 0.f;
        const int nav = transport_bound<1>(G, s_nav.layers, F.nav, x, c, sf, n_rk, leg, 0u,
                                           &s_nav, F.n_groups);
        if (nav == NAV_EXIT) {
            status = kIncomplete;
            break;
        }
        if (nav != NAV_FOUND) {
            status = kPropagationFailure;
            break;
        }
    }
    F.fw_info[t] = status | (min(holes, 255u) << 8) | (n_seq << 16);
}

__global__ __launch_bounds__(kFitBlock) void k_fit_backward(fit_params F, uint32_t iteration) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= F.n_tracks) {
        return;
    }
    if (iteration > 0u && F.outcome[t] != CKF_OUTCOME_SUCCESS) {
        return;
    }
    const uint32_t info = F.fw_info[t];
    const uint32_t fw = info & 0xffu;
    const uint32_t holes = (info >> 8) & 0xffu;
    const uint32_t n_seq = info >> 16;
    F.out_nholes[t] = holes;
    F.out_ndf[t] = 0.f;
    F.out_chi2[t] = 0.f;
    F.out_pval[t] = 0.f;
    if (fw != kSuccess) {
        F.outcome[t] = fw == kUpdaterSkipped      ? CKF_OUTCOME_FAILURE_NOT_ALL_FITTED
                       : fw == kPropagationFailure ? CKF_OUTCOME_FAILURE_FORWARD_PROPAGATION
                       : (fw == kSequenceOverflow || fw == kIncomplete) ? CKF_OUTCOME_FAILURE_SMOOTHER
                                                                         : CKF_OUTCOME_FAILURE_FITTER;
        return;
    }
    const uint32_t s0 = F.offsets[t];
    const uint32_t n = F.offsets[t + 1u] - s0;
    const uint32_t moff = F.ev_meas_off[track_event(F, t)];
    const uint32_t* seq = F.seq + static_cast<size_t>(s0) * F.seq_factor + static_cast<size_t>(t) * F.seq_min;
    const geo_view G{F.sfs, F.n_layers, F.field};
    // the last state: smoothed = filtered, then its backward pass starts
    // from the inflated filtered state
    float x[6], c[21];
    load_rec(F.flt + static_cast<size_t>(s0 + n - 1u) * kRec, x, c);
    {
        float* r = F.smt + static_cast<size_t>(s0 + n - 1u) * kRec;
        store_rec(r, x, c);
    }
    inflate(c, F.inflation);
    uint32_t sf = F.meas_sf[moff + F.cons[s0 + n - 1u]];
    uint32_t j = n;  // states still to smooth: j - 1 is the next
    uint32_t status = kSuccess;
    for (uint32_t e = n_seq; e-- > 0u;) {
        const uint32_t entry = seq[e];
        const uint32_t target = entry & ~kSeqState;
        if (target != sf) {
            if (transport_to_surface(G, F.nav, x, c, sf, target) != 0) {
                status = kPropagationFailure;
                break;
            }
        }
        if (entry & kSeqState) {
            --j;
            const uint32_t g = s0 + j;
            const meas_data md = load_meas(F.meas_pos, F.meas_var, F.meas_info, moff + F.cons[g]);
            float xf[6], cf[21], xs[6], cs[21];
            load_rec(F.flt + static_cast<size_t>(g) * kRec, xf, cf);
            float chi2_s = 0.f, chi2_b = 0.f;
            status = smooth_state(md, x, c, xf, cf, xs, cs, chi2_s, chi2_b);
            if (status != kSuccess) {
                break;
            }
            float* r = F.smt + static_cast<size_t>(g) * kRec;
            store_rec(r, xs, cs);
            r[kRecChi2a] = chi2_s;
            r[kRecChi2b] = chi2_b;
            F.flags[g] = CKF_STATE_SMOOTHED;
        }
        add_material_sf(F.sfs[target], F.nav.mass, F.nav.abs_q, x, c);
    }
    if (status != kSuccess) {
        F.outcome[t] = status == kSmootherSkipped     ? CKF_OUTCOME_FAILURE_NOT_ALL_SMOOTHED
                       : status == kPropagationFailure ? CKF_OUTCOME_FAILURE_BACKWARD_PROPAGATION
                                                       : CKF_OUTCOME_FAILURE_SMOOTHER;
        return;
    }
    // update_statistics: sums over smoothed states, parameters of the first
    float ndf = 0.f, chi2 = 0.f;
    bool have = false;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t g = s0 + k;
        if (F.flags[g] & CKF_STATE_SMOOTHED) {
            const float* r = F.smt + static_cast<size_t>(g) * kRec;
            if (!have) {
                float xs[6], cs[21];
                load_rec(r, xs, cs);
#pragma unroll
                for (int q = 0; q < 6; ++q) F.out_vec[6u * static_cast<size_t>(t) + q] = xs[q];
#pragma unroll
                for (int q = 0; q < 21; ++q) F.out_cov[21u * static_cast<size_t>(t) + q] = cs[q];
                F.out_surface[t] = F.meas_sf[moff + F.cons[g]];
                have = true;
            }
            ndf += static_cast<float>(F.meas_info[moff + F.cons[g]] & 0xffu);
            chi2 += r[kRecChi2b];
        }
    }
    ndf -= 5.f;
    F.out_ndf[t] = ndf;
    F.out_chi2[t] = chi2;
    F.out_pval[t] = prob_chi2(chi2, ndf);
    F.outcome[t] = ndf > 0.f ? CKF_OUTCOME_SUCCESS : CKF_OUTCOME_FAILURE_NON_POSITIVE_NDF;
}

}  // namespace
}  // namespace tamd

using tamd::ckf_error;
using tamd::guarded;
using tamd::hip_check;

struct ckf_fitter {
    ckf_fitting_config cfg{};
    hipStream_t stream = nullptr;
    tamd::dbuf seq, fw_info, ev_track_off;
};

namespace {

void validate(const ckf_fitting_config& c) {
    if (c.n_iterations == 0u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "n_iterations must be at least 1");
    }
    if (!(c.covariance_inflation_factor > 0.f) || c.surface_sequence_size_factor == 0u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "invalid smoother settings");
    }
    if (c.max_step_counts == 0u || !(std::fabs(c.ptc_charge) > 0.f) || !(c.ptc_mass >= 0.f)) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "invalid propagation settings");
    }
    if (c.propagation.search_window[0] > 1u || c.propagation.search_window[1] > 1u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "search_window is limited to 1 bin per side");
    }
    if (c.do_precise_hole_count != 0u) {
        throw ckf_error(CKF_ERR_UNSUPPORTED, "precise hole counting is not supported");
    }
}

tamd::nav_config<float> fit_nav(const ckf_fitting_config& c) {
    ckf_finding_config f{};
    ckf_finding_config_default(&f);
    f.propagation = c.propagation;
    f.min_step_length_for_next_surface = 0.f;  // no ckf_aborter in the fit
    f.max_step_counts_for_next_surface = c.max_step_counts;
    f.min_p = c.min_p;
    f.min_pT = c.min_pT;
    f.ptc_pdg = c.ptc_pdg;
    f.ptc_mass = c.ptc_mass;
    f.ptc_charge = c.ptc_charge;
    tamd::nav_config<float> n = tamd::make_nav_config<float>(f);
    // the fit reaches a module when its crossing lies within
    // min_mask_tolerance of the mask (the propagation configuration's mask
    // tolerance at zero distance); measurements on the edge of an overlap
    // module, which an exact mask test can miss by a few microns after the
    // refit, are then still reached
    n.accept_tol = static_cast<float>(c.propagation.min_mask_tolerance);
    return n;
}

template <typename T>
void d2h(T* dst, const void* src, size_t n, hipStream_t s) {
    if (dst != nullptr && n > 0u) {
        hip_check(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s), "download");
    }
}

}  // namespace

extern "C" {

ckf_status ckf_fitter_create(const ckf_fitting_config* cfg, ckf_fitter** out) {
    return guarded([&] {
        if (cfg == nullptr || out == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        validate(*cfg);
        auto* f = new ckf_fitter();
        f->cfg = *cfg;
        *out = f;
    });
}

ckf_status ckf_fitter_destroy(ckf_fitter* f) {
    return guarded([&] {
        if (f != nullptr) {
            f->seq.release();
            f->fw_info.release();
            f->ev_track_off.release();
        }
        delete f;
    });
}

ckf_status ckf_fitter_set_stream(ckf_fitter* f, void* hip_stream) {
    return guarded([&] {
        if (f == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null fitter");
        f->stream = static_cast<hipStream_t>(hip_stream);
    });
}

ckf_status ckf_fit_tracks(ckf_fitter* f, const ckf_device_detector* det, const ckf_device_event* ev,
                          const ckf_device_tracks* cand, ckf_device_fitted** out) {
    return guarded([&] {
        if (f == nullptr || det == nullptr || ev == nullptr || cand == nullptr || out == nullptr) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        if (ev->n_surfaces != det->h_sfs.size() || cand->n_events != ev->n_events) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "candidates, events and detector do not match");
        }
        const hipStream_t s = f->stream;
        const uint32_t n = cand->n_tracks, ns = cand->n_const;
        auto* R = new ckf_device_fitted();
        try {
            R->n_tracks = n;
            R->n_states = ns;
            R->ids = det->ids;
            R->ev_track_off = cand->ev_track_off;
            R->outcome.ensure(2u * std::max(n, 1u));
            R->surface.ensure(4u * std::max(n, 1u));
            R->vec.ensure(24u * std::max(n, 1u));
            R->cov.ensure(84u * std::max(n, 1u));
            R->ndf.ensure(4u * std::max(n, 1u));
            R->chi2.ensure(4u * std::max(n, 1u));
            R->pval.ensure(4u * std::max(n, 1u));
            R->nholes.ensure(4u * std::max(n, 1u));
            R->offsets.ensure(4u * (n + 1u));
            R->cons.ensure(4u * std::max(ns, 1u));
            R->flt.ensure(sizeof(float) * tamd::kRec * std::max<size_t>(ns, 1u));
            R->smt.ensure(sizeof(float) * tamd::kRec * std::max<size_t>(ns, 1u));
            R->flags.ensure(std::max(ns, 1u));
            R->state_sf.ensure(4u * std::max(ns, 1u));
            hip_check(hipMemsetAsync(R->flags.p, 0, R->flags.bytes, s), "memset");
            hip_check(hipMemsetAsync(R->state_sf.p, 0, R->state_sf.bytes, s), "memset");
            hip_check(hipMemsetAsync(R->outcome.p, 0, R->outcome.bytes, s), "memset");
            hip_check(hipMemsetAsync(R->surface.p, 0, R->surface.bytes, s), "memset");
            hip_check(hipMemsetAsync(R->vec.p, 0, R->vec.bytes, s), "memset");
            hip_check(hipMemsetAsync(R->cov.p, 0, R->cov.bytes, s), "memset");
            hip_check(hipMemsetAsync(R->smt.p, 0, R->smt.bytes, s), "memset");
            hip_check(hipMemsetAsync(R->flt.p, 0, R->flt.bytes, s), "memset");
            hip_check(hipMemcpyAsync(R->offsets.p, cand->offsets, 4u * (n + 1u),
                                     hipMemcpyDeviceToDevice, s), "copy offsets");
            if (ns > 0u) {
                hip_check(hipMemcpyAsync(R->cons.p, cand->cons, 4u * ns, hipMemcpyDeviceToDevice, s),
                          "copy constituents");
            }
            if (n > 0u) {
                // per-track surface sequences: n_states * factor + min per track
                const size_t seq_n = size_t(ns) * f->cfg.surface_sequence_size_factor +
                                     size_t(n) * f->cfg.min_surface_sequence_capacity;
                f->seq.ensure(4u * std::max<size_t>(seq_n, 1u));
                f->fw_info.ensure(4u * size_t(n));
                f->ev_track_off.ensure(4u * (cand->ev_track_off.size()));
                hip_check(hipMemcpyAsync(f->ev_track_off.p, cand->ev_track_off.data(),
                                         4u * cand->ev_track_off.size(), hipMemcpyHostToDevice, s),
                          "event offsets");
                tamd::fit_params F{};
                F.sfs = det->d_sfs;
                F.layers = det->d_layers;
                F.n_layers = static_cast<uint32_t>(det->h_layers.size());
                F.groups = det->d_groups;
                F.n_groups = static_cast<uint32_t>(det->h_groups.size());
                F.field = tamd::make_field_view<float>(det->field, det->d_map);
                F.nav = fit_nav(f->cfg);
                F.inflation = f->cfg.covariance_inflation_factor;
                F.seq_factor = f->cfg.surface_sequence_size_factor;
                F.seq_min = f->cfg.min_surface_sequence_capacity;
                F.meas_pos = ev->d_pos;
                F.meas_var = ev->d_var;
                F.meas_info = ev->d_info;
                F.meas_sf = ev->d_meas_sf;
                F.ev_meas_off = ev->d_meas_off;
                F.ev_track_off = f->ev_track_off.as<uint32_t>();
                F.n_events = ev->n_events;
                F.n_tracks = n;
                F.cand_surface = cand->surface;
                F.cand_vec = cand->vec;
                F.cand_cov = cand->cov;
                F.offsets = R->offsets.as<uint32_t>();
                F.cons = R->cons.as<uint32_t>();
                F.seq = f->seq.as<uint32_t>();
                F.fw_info = f->fw_info.as<uint32_t>();
                F.flt = R->flt.as<float>();
                F.smt = R->smt.as<float>();
                F.flags = R->flags.as<uint8_t>();
                F.state_sf = R->state_sf.as<uint32_t>();
                F.outcome = R->outcome.as<uint16_t>();
                F.out_surface = R->surface.as<uint32_t>();
                F.out_vec = R->vec.as<float>();
                F.out_cov = R->cov.as<float>();
                F.out_ndf = R->ndf.as<float>();
                F.out_chi2 = R->chi2.as<float>();
                F.out_pval = R->pval.as<float>();
                F.out_nholes = R->nholes.as<uint32_t>();
                const uint32_t blocks = (n + tamd::kFitBlock - 1u) / tamd::kFitBlock;
                for (uint32_t it = 0; it < f->cfg.n_iterations; ++it) {
                    hipLaunchKernelGGL(tamd::k_fit_forward, dim3(blocks), dim3(tamd::kFitBlock), 0, s, F, it);
                    hip_check(hipGetLastError(), "fit forward");
                    hipLaunchKernelGGL(tamd::k_fit_backward, dim3(blocks), dim3(tamd::kFitBlock), 0, s, F, it);
                    hip_check(hipGetLastError(), "fit backward");
                }
            }
            hip_check(hipStreamSynchronize(s), "fit");
        } catch (...) {
            delete R;
            throw;
        }
        *out = R;
    });
}

ckf_status ckf_fitted_size(const ckf_device_fitted* r, uint32_t* n_tracks, uint32_t* n_states) {
    return guarded([&] {
        if (r == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null result");
        if (n_tracks) *n_tracks = r->n_tracks;
        if (n_states) *n_states = r->n_states;
    });
}

ckf_status ckf_fitted_download(const ckf_device_fitted* r, ckf_fitted_tracks* o) {
    return guarded([&] {
        if (r == nullptr || o == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        if (o->n_tracks != r->n_tracks || o->n_states != r->n_states) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "output sized for another result");
        }
        const uint32_t n = r->n_tracks, ns = r->n_states;
        const hipStream_t s = nullptr;
        std::vector<uint32_t> sfs(n), cons(ns), state_sf(ns);
        std::vector<float> cov21(21u * size_t(n)), flt(size_t(ns) * tamd::kRec),
            smt(size_t(ns) * tamd::kRec);
        d2h(o->fit_outcome, r->outcome.p, n, s);
        d2h(sfs.data(), r->surface.p, n, s);
        d2h(o->params_vector, r->vec.p, 6u * size_t(n), s);
        d2h(cov21.data(), r->cov.p, 21u * size_t(n), s);
        d2h(o->ndf, r->ndf.p, n, s);
        d2h(o->chi2, r->chi2.p, n, s);
        d2h(o->pval, r->pval.p, n, s);
        d2h(o->nholes, r->nholes.p, n, s);
        d2h(o->state_offsets, r->offsets.p, n + 1u, s);
        d2h(o->state_flags, r->flags.p, ns, s);
        d2h(cons.data(), r->cons.p, ns, s);
        d2h(state_sf.data(), r->state_sf.p, ns, s);
        d2h(flt.data(), r->flt.p, flt.size(), s);
        d2h(smt.data(), r->smt.p, smt.size(), s);
        hip_check(hipStreamSynchronize(s), "download");
        auto unpack = [](const float* c21, float* c36) {
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) c36[6 * i + j] = c21[tamd::sym_idx(i, j)];
        };
        for (uint32_t t = 0; t < n; ++t) {
            if (o->params_surface_link) {
                const uint16_t oc = o->fit_outcome ? o->fit_outcome[t] : CKF_OUTCOME_SUCCESS;
                o->params_surface_link[t] = (oc == CKF_OUTCOME_SUCCESS && sfs[t] < r->ids.size())
                                                ? r->ids[sfs[t]] : ~0ull;
            }
            if (o->params_covariance) unpack(&cov21[21u * size_t(t)], o->params_covariance + 36u * size_t(t));
        }
        for (uint32_t g = 0; g < ns; ++g) {
            const float* fr = &flt[size_t(g) * tamd::kRec];
            const float* sr = &smt[size_t(g) * tamd::kRec];
            if (o->filtered_chi2) o->filtered_chi2[g] = fr[tamd::kRecChi2a];
            if (o->smoothed_chi2) o->smoothed_chi2[g] = sr[tamd::kRecChi2a];
            if (o->backward_chi2) o->backward_chi2[g] = sr[tamd::kRecChi2b];
            if (o->filtered_vector) std::memcpy(o->filtered_vector + 6u * size_t(g), fr, 24u);
            if (o->smoothed_vector) std::memcpy(o->smoothed_vector + 6u * size_t(g), sr, 24u);
            if (o->filtered_covariance) unpack(fr + 6, o->filtered_covariance + 36u * size_t(g));
            if (o->smoothed_covariance) unpack(sr + 6, o->smoothed_covariance + 36u * size_t(g));
            if (o->measurement_index) o->measurement_index[g] = cons[g];
            if (o->state_surface_link) {
                o->state_surface_link[g] = state_sf[g] < r->ids.size() ? r->ids[state_sf[g]] : ~0ull;
            }
        }
    });
}

ckf_status ckf_fitted_destroy(ckf_device_fitted* r) {
    return guarded([&] { delete r; });
}

}  // extern "C"
