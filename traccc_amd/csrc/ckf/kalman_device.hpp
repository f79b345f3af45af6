// FP32 Kalman pieces on the symmetric-packed bound state: predicted chi2 and
// the gain-matrix update with a Joseph-form covariance.
//
// Restates, for one (branch, measurement) pair and in registers:
//   measurement_selector::predicted_chi2  core/include/traccc/finding/measurement_selector.hpp:152-212
//   gain_matrix_updater::operator()       core/include/traccc/fitting/kalman_filter/gain_matrix_updater.hpp:69-181
//   normalize_angles                      core/include/traccc/edm/track_parameters.hpp:52-83
//   regularize_covariance                 core/include/traccc/fitting/details/regularize_covariance.hpp:20-50
// The projector H only selects (and for line surfaces negates) bound
// components, so H P H^T, P H^T and the Joseph product are formed from
// columns of P instead of dense 6x6 products.
#pragma once

#include <cmath>
#include <cfloat>

#include "device_types.hpp"

namespace tamd {

struct meas_data {
    float m0, m1;   // local position in subspace order
    float v0, v1;   // variance in subspace order; v1 = FLT_MAX for 1-D
    uint32_t dims;
    uint32_t a, b;  // bound indices measured
};

TAMD_HD meas_data load_meas(const float2* pos, const float2* var, const uint32_t* info,
                            uint32_t i) {
    meas_data m;
    const float2 p = pos[i];
    const float2 v = var[i];
    const uint32_t inf = info[i];
    m.dims = inf & 0xffu;
    m.a = (inf >> 8) & 0xffu;
    m.b = (inf >> 16) & 0xffu;
    // get_measurement_local: a subspace starting at loc1 swaps the coordinates
    if (m.a == 1u) {
        m.m0 = p.y; m.m1 = p.x; m.v0 = v.y; m.v1 = v.x;
    } else {
        m.m0 = p.x; m.m1 = p.y; m.v0 = v.x; m.v1 = v.y;
    }
    if (m.dims == 1u) {
        m.v1 = FLT_MAX;
        m.m1 = 0.f;
    }
    return m;
}

// Measured bound indices are loc0 / loc1 (a, b in {0, 1}; the library
// rejects other subspaces at upload), so every runtime-indexed access below is
// a two-way select between registers.
TAMD_HD float pick01(const float* x, uint32_t k) { return k ? x[1] : x[0]; }

// Column k (k in {0, 1}) of the packed covariance.
TAMD_HD void cov_col(const float* c, uint32_t k, float* col) {
TAMD_UNROLL
    for (int i = 0; i < 6; ++i) {
        col[i] = k ? c[sym_idx(i, 1)] : c[sym_idx(i, 0)];
    }
}

// The 2x2 block and sub-vector a measurement with subspace (a, b) sees.
struct proj_view {
    float x0, x1;       // x[a], x[b]
    float p00, p01, p11;
};

TAMD_HD proj_view project(const float* x, const float* c, uint32_t a, uint32_t b) {
    proj_view v;
    const float c00 = c[sym_idx(0, 0)], c01 = c[sym_idx(0, 1)], c11 = c[sym_idx(1, 1)];
    v.x0 = pick01(x, a);
    v.x1 = pick01(x, b);
    v.p00 = a ? c11 : c00;
    v.p11 = b ? c11 : c00;
    v.p01 = (a == b) ? v.p00 : c01;
    return v;
}

TAMD_HD float predicted_chi2(const meas_data& m, const proj_view& v) {
    const float r0 = m.m0 - v.x0;
    const float R00 = v.p00 + m.v0;
    if (m.dims == 1u) {
        return r0 * (1.f / R00) * r0;
    }
    const float r1 = m.m1 - v.x1;
    const float R01 = v.p01;
    const float R11 = v.p11 + m.v1;
    const float det = R00 * R11 - R01 * R01;
    const float i00 = R11 / det, i11 = R00 / det, i01 = -R01 / det;
    return r0 * (i00 * r0 + i01 * r1) + r1 * (i01 * r0 + i11 * r1);
}

TAMD_HD float predicted_chi2(const meas_data& m, const float* x, const float* c) {
    const proj_view v = project(x, c, m.a, m.b);
    return predicted_chi2(m, v);
}

// wrap_phi_theta semantics (core/include/traccc/utils/trigonometric_helpers.hpp:63-81)
TAMD_HD bool normalize_angles(float& phi, float& theta) {
    const float pi = 3.14159265358979323846f;
    const float twopi = 6.28318530717958647692f;
    const float in_theta = theta;
    if (std::fmod(in_theta, twopi) == 0.f || in_theta == 0.f) {
        return false;
    }
    float th = in_theta;
    if (!(th >= 0.f && th < twopi)) {
        th = th - twopi * std::floor(th / twopi);
    }
    float ph = phi;
    if (pi < th) {
        ph += pi;
        th = twopi - th;
    }
    const float d = ph + pi;
    if (!(d >= 0.f && d < twopi)) {
        ph = ph - twopi * std::floor(d / twopi);
    }
    if (th <= 0.f || th >= twopi) {
        return false;
    }
    phi = ph;
    theta = th;
    return true;
}

// Gain-matrix update in place for measured bound indices (a, b): x (6),
// c (21 packed).  Returns a ckf_fitter_status; x and c are left untouched on
// failure.
#if defined(__HIPCC__)
__attribute__((always_inline))
#endif
TAMD_HD int gain_update_ab(const meas_data& m, uint32_t a, uint32_t b, float* x, float* c) {
    float pa[6], pb[6];  // columns a and b of P = P H^T
    cov_col(c, a, pa);
    cov_col(c, b, pb);
    const float paa = pick01(pa, a), pab = pick01(pa, b), pba = pick01(pb, a), pbb = pick01(pb, b);
    const float R00 = paa + m.v0;
    float k0[6], k1[6];
    float r0 = m.m0 - pick01(x, a);
    float r1 = 0.f;
    if (m.dims == 1u) {
        const float inv = 1.f / R00;
        TAMD_UNROLL
        for (int i = 0; i < 6; ++i) {
            k0[i] = pa[i] * inv;
            k1[i] = 0.f;
        }
    } else {
        const float R01 = pab;
        const float R11 = pbb + m.v1;
        const float det = R00 * R11 - R01 * R01;
        const float i00 = R11 / det, i11 = R00 / det, i01 = -R01 / det;
        TAMD_UNROLL
        for (int i = 0; i < 6; ++i) {
            k0[i] = pa[i] * i00 + pb[i] * i01;
            k1[i] = pa[i] * i01 + pb[i] * i11;
        }
        r1 = m.m1 - pick01(x, b);
    }
    float nx[6];
    TAMD_UNROLL
    for (int i = 0; i < 6; ++i) {
        nx[i] = x[i] + (k0[i] * r0 + k1[i] * r1);
    }
    if (!std::isfinite(nx[e_theta]) || !std::isfinite(nx[e_phi])) {
        return CKF_FIT_ERROR_INVERSION;
    }
    if (std::fabs(nx[e_qop]) == 0.f) {
        return CKF_FIT_ERROR_QOP_ZERO;
    }
    if (!normalize_angles(nx[e_phi], nx[e_theta])) {
        return CKF_FIT_ERROR_THETA_POLE;
    }
    // Joseph form: P' = (I - K H) P (I - K H)^T + K V K^T.
    // Q = (I - K H) P has columns Q[:, j] = P[:, j] - k0 P[a, j] - k1 P[b, j].
    float qa[6], qb[6];
    TAMD_UNROLL
    for (int i = 0; i < 6; ++i) {
        qa[i] = pa[i] - k0[i] * paa - k1[i] * pba;
        qb[i] = pb[i] - k0[i] * pab - k1[i] * pbb;
    }
    const bool two = (m.dims == 2u);
    float nc[21];
    TAMD_UNROLL
    for (int i = 0; i < 6; ++i) {
        TAMD_UNROLL
        for (int j = i; j < 6; ++j) {
            const float qij = c[sym_idx(i, j)] - k0[i] * pa[j] - k1[i] * pb[j];
            float v = qij - qa[i] * k0[j] - qb[i] * k1[j] + k0[i] * m.v0 * k0[j];
            if (two) {
                v += k1[i] * m.v1 * k1[j];
            }
            nc[sym_idx(i, j)] = v;
        }
    }
    const float min_var = -0.01f;
    TAMD_UNROLL
    for (int i = 0; i < 6; ++i) {
        if (nc[sym_idx(i, i)] < min_var) {
            return CKF_FIT_ERROR_UPDATER_INVALID_COVARIANCE;
        }
    }
    TAMD_UNROLL
    for (int i = 0; i < 6; ++i) {
        nc[sym_idx(i, i)] = std::fabs(nc[sym_idx(i, i)]);
    }
    TAMD_UNROLL
    for (int i = 0; i < 6; ++i) {
        x[i] = nx[i];
    }
    TAMD_UNROLL
    for (int i = 0; i < 21; ++i) {
        c[i] = nc[i];
    }
    return CKF_FIT_SUCCESS;
}

// The (loc0, loc1) pixel measurement gets its own copy with the subspace as
// constants, which folds every column / element select away; any other
// subspace takes the general copy.
TAMD_HD int gain_update(const meas_data& m, float* x, float* c) {
    if (m.a == 0u && m.b == 1u) {
        return gain_update_ab(m, 0u, 1u, x, c);
    }
    return gain_update_ab(m, m.a, m.b, x, c);
}

// chi2 p-value Q(ndf / 2, chi2 / 2): the reference's split between a power
// series (x < 1 or x < a) and a continued fraction (core/include/traccc/utils/prob.hpp:148-288),
// restated in FP32; log-gamma from the device math library.
TAMD_HD float prob_chi2(float chi2, float ndf) {
    if (ndf <= 0.f) {
        return 0.f;
    }
    if (chi2 <= 0.f) {
        return chi2 < 0.f ? 0.f : 1.f;
    }
    const float a = 0.5f * ndf;
    const float x = 0.5f * chi2;
    const float eps = 1.11022302462515654042e-16f;
    float ax = a * std::log(x) - x - std::lgamma(a);
    if (ax < -709.78f) {
        return (x < 1.f || x < a) ? 1.f : 0.f;
    }
    ax = std::exp(ax);
    if (x < 1.f || x < a) {
        float r = a, cc = 1.f, ans = 1.f;
        int guard = 0;
        do {
            r += 1.f;
            cc *= x / r;
            ans += cc;
        } while (cc / ans > eps && ++guard < 2000);
        return 1.f - ans * ax / a;
    }
    float y = 1.f - a, z = x + y + 1.f, cc = 0.f;
    float pkm2 = 1.f, qkm2 = x, pkm1 = x + 1.f, qkm1 = z * x;
    float ans = pkm1 / qkm1, t = 1.f;
    int guard = 0;
    do {
        cc += 1.f;
        y += 1.f;
        z += 2.f;
        const float yc = y * cc;
        const float pk = pkm1 * z - pkm2 * yc;
        const float qk = qkm1 * z - qkm2 * yc;
        if (qk != 0.f) {
            const float r = pk / qk;
            t = std::fabs((ans - r) / r);
            ans = r;
        } else {
            t = 1.f;
        }
        pkm2 = pkm1;
        pkm1 = pk;
        qkm2 = qkm1;
        qkm1 = qk;
        if (std::fabs(pk) > 4.503599627370496e15f) {
            pkm2 *= 2.22044604925031308085e-16f;
            pkm1 *= 2.22044604925031308085e-16f;
            qkm2 *= 2.22044604925031308085e-16f;
            qkm1 *= 2.22044604925031308085e-16f;
        }
    } while (t > eps && ++guard < 2000);
    return ans * ax;
}

}  // namespace tamd
