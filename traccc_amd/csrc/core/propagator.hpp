// Track propagation to the next sensitive surface: RK4-Nystroem stepping with
// covariance (Jacobian) transport, and a layer-window navigator.
//
// What this replaces: the CKF's detray propagator
//   rk_stepper + caching_navigator + {pathlimit_aborter, parameter_updater,
//   momentum_aborter, ckf_aborter}
// (core/include/traccc/finding/details/combinatorial_kalman_filter_types.hpp:27-52,
//  core/include/traccc/finding/actors/ckf_aborter.hpp:37-68), called from
// core/include/traccc/finding/details/combinatorial_kalman_filter.hpp:528-627.
//
// Methods (published, restated here in this project's own code):
//  * Runge-Kutta-Nystroem 4th order for r'' = (q/p) r' x B
//    (Abramowitz & Stegun 25.5.20; Lund et al., JINST 4 (2009) P04001);
//  * semi-analytic transport of d(free)/d(bound) through the RK stages
//    (Bugge & Myrheim, NIM 179 (1981) 365), field gradients neglected;
//  * path-length correction onto the destination plane, then the
//    free-to-bound projection.
//
// Step size: a-priori accuracy control.  For a helix of curvature kappa the
// RK4 local error is about h^5 kappa^4 / 120, so the step is
//   h = (120 * rk_error_tol)^(1/5) * kappa^(-4/5), capped by step_constraint.
// It is a smooth function of the state, so an FP32 and an FP64 evaluation of
// the same track take the same sequence of steps.
#pragma once

#include <cmath>
#include <type_traits>

#include "defs.hpp"
#include "field.hpp"
#include "geometry.hpp"

namespace tamd {

// Position carried as an unevaluated sum pos + lo (Knuth TwoSum on every
// increment): in FP32 the rounding of |x| ~ 1 m positions would otherwise
// reach 1e-4 mm, a visible fraction of a 50 um residual.
template <typename T>
struct free_state {
    vec3<T> pos;
    vec3<T> dir;
    T qop;
    T time;
    vec3<T> lo;
};

template <typename T>
TAMD_HD void two_sum_add(T& hi, T& lo, T d) {
    const T s = hi + d;
    const T bp = s - hi;
    const T err = (hi - (s - bp)) + (d - bp);
    hi = s;
    lo += err;
}

// (pos + lo) - c, evaluated without losing lo
template <typename T>
TAMD_HD vec3<T> rel_to(const free_state<T>& s, const vec3<T>& c) {
    return {(s.pos.x - c.x) + s.lo.x, (s.pos.y - c.y) + s.lo.y, (s.pos.z - c.z) + s.lo.z};
}

// d(free)/d(bound at the start surface); bound columns: loc0, loc1, phi,
// theta, q/p (the time column only feeds the time row and is implicit).
template <typename T>
struct transport_jac {
    T p[3][5];
    T d[3][5];
    T t_qop;
};

template <typename T>
struct nav_config {
    T min_step_length;   // ckf_aborter: path needed before a surface counts
    uint32_t max_steps;  // ckf_aborter: RK steps allowed per propagation
    T on_surface_tol;
    T overstep_tol;
    T mask_tol_min, mask_tol_max, mask_tol_scalor;
    T step_constraint;
    T rk_step_scale;  // (120 * rk_error_tol)^(1/5)
    T path_limit;
    T approach_dist;
    int win_phi, win_long;
    T mass;   // particle hypothesis mass, GeV
    T abs_q;  // |charge| of the hypothesis
    T min_p, min_pT;
    // a module is reached when the crossing lies within this distance of its
    // mask (0, the CKF: the exact mask)
    T accept_tol;
};

template <typename T>
inline nav_config<T> make_nav_config(const ckf_finding_config& c) {
    nav_config<T> n{};
    n.min_step_length = static_cast<T>(c.min_step_length_for_next_surface);
    n.max_steps = c.max_step_counts_for_next_surface;
    n.on_surface_tol = static_cast<T>(c.propagation.on_surface_tolerance);
    n.overstep_tol = static_cast<T>(c.propagation.overstep_tolerance);
    n.mask_tol_min = static_cast<T>(c.propagation.min_mask_tolerance);
    n.mask_tol_max = static_cast<T>(c.propagation.max_mask_tolerance);
    n.mask_tol_scalor = static_cast<T>(c.propagation.mask_tolerance_scalor);
    n.step_constraint = static_cast<T>(c.propagation.step_constraint);
    n.rk_step_scale = static_cast<T>(
        std::pow(120.0 * static_cast<double>(c.propagation.rk_error_tol), 0.2));
    n.path_limit = static_cast<T>(c.propagation.path_limit);
    n.approach_dist = static_cast<T>(25.0);
    n.win_phi = static_cast<int>(c.propagation.search_window[0]);
    n.win_long = static_cast<int>(c.propagation.search_window[1]);
    n.mass = static_cast<T>(c.ptc_mass);
    n.abs_q = static_cast<T>(std::fabs(c.ptc_charge));
    n.min_p = static_cast<T>(c.min_p);
    n.min_pT = static_cast<T>(c.min_pT);
    n.accept_tol = T(0);
    return n;
}

// dt/ds = 1 / (beta c) for the hypothesis mass.
template <typename T>
TAMD_HD T dtds_of(T qop, T mass, T abs_q) {
    const T mq = mass * qop / abs_q;
    return std::sqrt(T(1) + mq * mq) / static_cast<T>(kSpeedOfLight);
}

// One RK4-Nystroem step of length h (may be negative).
template <typename T, bool kJac>
TAMD_HD void rk4_step(free_state<T>& s, transport_jac<T>& J,
                      const field_view<T>& f, T h, T mass, T abs_q) {
    const T q = s.qop;
    const T h2 = h * T(0.5);
    const vec3<T> B1 = field_at(f, s.pos);
    const vec3<T> k1 = q * cross(s.dir, B1);
    const vec3<T> p2 = s.pos + h2 * s.dir + (h * h * T(0.125)) * k1;
    const vec3<T> B2 = field_at(f, p2);
    const vec3<T> T2 = s.dir + h2 * k1;
    const vec3<T> k2 = q * cross(T2, B2);
    const vec3<T> T3 = s.dir + h2 * k2;
    const vec3<T> k3 = q * cross(T3, B2);
    const vec3<T> p4 = s.pos + h * s.dir + (h * h * T(0.5)) * k3;
    const vec3<T> B4 = field_at(f, p4);
    const vec3<T> T4 = s.dir + h * k3;
    const vec3<T> k4 = q * cross(T4, B4);

    const T h6 = h / T(6);
    const T hh6 = h * h / T(6);

    if constexpr (kJac) {
        // Columns of dk_i/dT0 (3x3) and dk_i/dq.
        T A1[3][3], A2[3][3], A3[3][3], A4[3][3];
        TAMD_UNROLL
        for (int c = 0; c < 3; ++c) {
            const vec3<T> e{c == 0 ? T(1) : T(0), c == 1 ? T(1) : T(0),
                            c == 2 ? T(1) : T(0)};
            const vec3<T> a1 = q * cross(e, B1);
            const vec3<T> a2 = q * cross(e + h2 * a1, B2);
            const vec3<T> a3 = q * cross(e + h2 * a2, B2);
            const vec3<T> a4 = q * cross(e + h * a3, B4);
            A1[0][c] = a1.x; A1[1][c] = a1.y; A1[2][c] = a1.z;
            A2[0][c] = a2.x; A2[1][c] = a2.y; A2[2][c] = a2.z;
            A3[0][c] = a3.x; A3[1][c] = a3.y; A3[2][c] = a3.z;
            A4[0][c] = a4.x; A4[1][c] = a4.y; A4[2][c] = a4.z;
        }
        const vec3<T> d1 = cross(s.dir, B1);
        const vec3<T> d2 = cross(T2, B2) + q * cross(h2 * d1, B2);
        const vec3<T> d3 = cross(T3, B2) + q * cross(h2 * d2, B2);
        const vec3<T> d4 = cross(T4, B4) + q * cross(h * d3, B4);
        const T drl[3] = {hh6 * (d1.x + d2.x + d3.x), hh6 * (d1.y + d2.y + d3.y),
                          hh6 * (d1.z + d2.z + d3.z)};
        const T dtl[3] = {h6 * (d1.x + T(2) * d2.x + T(2) * d3.x + d4.x),
                          h6 * (d1.y + T(2) * d2.y + T(2) * d3.y + d4.y),
                          h6 * (d1.z + T(2) * d2.z + T(2) * d3.z + d4.z)};
        T DrT[3][3], DTT[3][3];
        TAMD_UNROLL
        for (int r = 0; r < 3; ++r) {
            TAMD_UNROLL
            for (int c = 0; c < 3; ++c) {
                const T id = (r == c) ? T(1) : T(0);
                DrT[r][c] = h * id + hh6 * (A1[r][c] + A2[r][c] + A3[r][c]);
                DTT[r][c] = id + h6 * (A1[r][c] + T(2) * A2[r][c] +
                                       T(2) * A3[r][c] + A4[r][c]);
            }
        }
        TAMD_UNROLL
        for (int j = 0; j < 5; ++j) {
            const T d0 = J.d[0][j], dd1 = J.d[1][j], dd2 = J.d[2][j];
            TAMD_UNROLL
            for (int r = 0; r < 3; ++r) {
                J.p[r][j] += DrT[r][0] * d0 + DrT[r][1] * dd1 + DrT[r][2] * dd2;
            }
            T nd[3];
            TAMD_UNROLL
            for (int r = 0; r < 3; ++r) {
                nd[r] = DTT[r][0] * d0 + DTT[r][1] * dd1 + DTT[r][2] * dd2;
            }
            J.d[0][j] = nd[0];
            J.d[1][j] = nd[1];
            J.d[2][j] = nd[2];
        }
        TAMD_UNROLL
        for (int r = 0; r < 3; ++r) {
            J.p[r][4] += drl[r];
            J.d[r][4] += dtl[r];
        }
        const T dt = dtds_of(q, mass, abs_q);
        const T mq = mass / abs_q;
        J.t_qop += h * (mq * mq * q) /
                   (static_cast<T>(kSpeedOfLight) * static_cast<T>(kSpeedOfLight) * dt);
    }

    const vec3<T> dp = h * s.dir + hh6 * (k1 + k2 + k3);
    two_sum_add(s.pos.x, s.lo.x, dp.x);
    two_sum_add(s.pos.y, s.lo.y, dp.y);
    two_sum_add(s.pos.z, s.lo.z, dp.z);
    vec3<T> nd = s.dir + h6 * (k1 + T(2) * k2 + T(2) * k3 + k4);
    const T inv = T(1) / std::sqrt(dot(nd, nd));
    s.dir = inv * nd;
    s.time += h * dtds_of(q, mass, abs_q);
}

// Bound parameters on surface sf -> free state; Jacobian initialised to
// d(free)/d(bound).
template <typename T, bool kJac>
TAMD_HD void bound_to_free(const ckf_surface& sf, const T* v, free_state<T>& s,
                           transport_jac<T>& J) {
    const vec3<T> c = sf_center<T>(sf);
    const vec3<T> u = sf_u<T>(sf);
    const vec3<T> w = sf_v<T>(sf);
    const vec3<T> d = v[e_loc0] * u + v[e_loc1] * w;
    s.pos = c;
    s.lo = {T(0), T(0), T(0)};
    two_sum_add(s.pos.x, s.lo.x, d.x);
    two_sum_add(s.pos.y, s.lo.y, d.y);
    two_sum_add(s.pos.z, s.lo.z, d.z);
    const T sp = std::sin(v[e_phi]), cp = std::cos(v[e_phi]);
    const T st = std::sin(v[e_theta]), ct = std::cos(v[e_theta]);
    s.dir = {cp * st, sp * st, ct};
    s.qop = v[e_qop];
    s.time = v[e_time];
    if constexpr (kJac) {
        TAMD_UNROLL
        for (int r = 0; r < 3; ++r) {
            TAMD_UNROLL
            for (int j = 0; j < 5; ++j) {
                J.p[r][j] = T(0);
                J.d[r][j] = T(0);
            }
        }
        J.p[0][0] = u.x; J.p[1][0] = u.y; J.p[2][0] = u.z;
        J.p[0][1] = w.x; J.p[1][1] = w.y; J.p[2][1] = w.z;
        J.d[0][2] = -sp * st; J.d[1][2] = cp * st; J.d[2][2] = T(0);
        J.d[0][3] = cp * ct;  J.d[1][3] = sp * ct; J.d[2][3] = -st;
        J.t_qop = T(0);
    }
}

// Free state on surface sf -> bound vector; with the path correction onto
// the plane, the full 6x6 bound-to-bound Jacobian.
template <typename T, bool kJac>
TAMD_HD void free_to_bound(const ckf_surface& sf, const free_state<T>& s,
                           const transport_jac<T>& J, const field_view<T>& f,
                           T mass, T abs_q, T* v, T (*Jb)[6]) {
    const vec3<T> c = sf_center<T>(sf);
    const vec3<T> u = sf_u<T>(sf);
    const vec3<T> vv = sf_v<T>(sf);
    const vec3<T> rel = rel_to(s, c);
    v[e_loc0] = dot(rel, u);
    v[e_loc1] = dot(rel, vv);
    v[e_phi] = std::atan2(s.dir.y, s.dir.x);
    T cz = s.dir.z;
    cz = cz > T(1) ? T(1) : (cz < T(-1) ? T(-1) : cz);
    v[e_theta] = std::acos(cz);
    v[e_qop] = s.qop;
    v[e_time] = s.time;
    if constexpr (kJac) {
        const vec3<T> w = sf_w<T>(sf);
        const T tw = dot(s.dir, w);
        const vec3<T> B = field_at(f, s.pos);
        const vec3<T> dTds = s.qop * cross(s.dir, B);
        const T dt = dtds_of(s.qop, mass, abs_q);
        T pc[3][5], dc[3][5], tc[5];
        TAMD_UNROLL
        for (int j = 0; j < 5; ++j) {
            const T dds = -(w.x * J.p[0][j] + w.y * J.p[1][j] + w.z * J.p[2][j]) / tw;
            pc[0][j] = J.p[0][j] + s.dir.x * dds;
            pc[1][j] = J.p[1][j] + s.dir.y * dds;
            pc[2][j] = J.p[2][j] + s.dir.z * dds;
            dc[0][j] = J.d[0][j] + dTds.x * dds;
            dc[1][j] = J.d[1][j] + dTds.y * dds;
            dc[2][j] = J.d[2][j] + dTds.z * dds;
            tc[j] = (j == 4 ? J.t_qop : T(0)) + dt * dds;
        }
        const T sin2 = s.dir.x * s.dir.x + s.dir.y * s.dir.y;
        const T sinT = std::sqrt(sin2);
        const vec3<T> gphi{-s.dir.y / sin2, s.dir.x / sin2, T(0)};
        const vec3<T> gth{s.dir.x * s.dir.z / sinT, s.dir.y * s.dir.z / sinT, -sinT};
        TAMD_UNROLL
        for (int j = 0; j < 5; ++j) {
            Jb[0][j] = u.x * pc[0][j] + u.y * pc[1][j] + u.z * pc[2][j];
            Jb[1][j] = vv.x * pc[0][j] + vv.y * pc[1][j] + vv.z * pc[2][j];
            Jb[2][j] = gphi.x * dc[0][j] + gphi.y * dc[1][j];
            Jb[3][j] = gth.x * dc[0][j] + gth.y * dc[1][j] + gth.z * dc[2][j];
            Jb[4][j] = (j == 4) ? T(1) : T(0);
            Jb[5][j] = tc[j];
        }
        TAMD_UNROLL
        for (int r = 0; r < 5; ++r) {
            Jb[r][5] = T(0);
        }
        Jb[5][5] = T(1);
    }
}

// Highland multiple-scattering width (PDG review, passage of particles
// through matter, eq. 34.15): theta0 = 13.6 MeV / (beta c p) |z| sqrt(x/X0)
// (1 + 0.038 ln(x z^2 / (X0 beta^2))).
template <typename T>
TAMD_HD T highland_theta0(T x_over_x0, T qop, T mass, T abs_q) {
    if (!(x_over_x0 > T(0))) {
        return T(0);
    }
    const T p = abs_q / std::fabs(qop);
    const T e = std::sqrt(p * p + mass * mass);
    const T beta = p / e;
    const T corr = T(1) + T(0.038) * std::log(x_over_x0 * abs_q * abs_q / (beta * beta));
    return T(0.0136) / (beta * p) * abs_q * std::sqrt(x_over_x0) * corr;
}

// NAV_DEFER: propagate(..., defer) stopped before its second layer search;
// layer_loop(deferred_visited, 1, ...) continues it.
enum nav_status : int { NAV_FOUND = 0, NAV_EXIT = 1, NAV_ABORT = 2, NAV_DEFER = 3 };

constexpr int kMaxCandidates = 9;

// Navigator + stepper for one track.  The algorithm:
//  1. If the start surface is given, visit the remaining modules of its layer
//     (module overlaps), then mark the layer visited.
//  2. Pick the unvisited layer whose (r, z) envelope the straight line from
//     the current state enters first; none -> NAV_EXIT.
//  3. Approach it with RK steps until the entry is within approach_dist.
//  4. Collect the modules in the search window around the projected entry
//     point (every module, up to 9, of a plane layer) whose straight-line
//     intersection is ahead (>= overstep_tol) and
//     inside the mask within a distance-scaled tolerance; visit them in
//     increasing straight-line distance: step onto each plane (iterated RK
//     steps until |distance| <= on_surface_tol, then one last step of that
//     distance), and
//     accept the first whose exact mask contains the point and whose path
//     from the start exceeds min_step_length -> NAV_FOUND.
//  5. Mark the layer visited, go to 2.
// More than max_steps RK steps, a path beyond path_limit, or a start momentum
// below min_p / min_pT -> NAV_ABORT.
//
// Team form (kT > 1, device code): kT consecutive lanes of a wave navigate one
// track together.  They hold the same state and take the same decisions; the
// layer search (step 2) and the window candidates (step 4) are split over the
// lanes and joined with shuffle reductions that keep the serial order's
// tie-breaks (lowest layer index, lowest module index), so every lane of the
// team reaches the serial navigator's result.  tl is the lane's index in its
// team.
template <typename T, bool kJac, int kT = 1>
struct navigator {
    static_assert(kT >= 1 && kT <= 64 && (kT & (kT - 1)) == 0, "team size: a power of two");
    static constexpr int kSlotsPerLane = (kMaxCandidates + kT - 1) / kT;

    const ckf_surface* sfs;
    const ckf_layer* layers;
    uint32_t n_layers;
    field_view<T> field;
    nav_config<T> cfg;
    uint32_t tl = 0u;
    // layer groups (build_layer_groups) and the boxes of the layers and the
    // groups (nav_box_of, this configuration's margin) for the one-lane layer
    // search; none: every layer is scanned
    const layer_group* groups = nullptr;
    uint32_t n_groups = 0u;
    const nav_box* layer_boxes = nullptr;
    const nav_box* group_boxes = nullptr;  // kMaxGroups slots

    free_state<T> st;
    transport_jac<T> J;
    T path;
    uint32_t n_steps;
    uint64_t deferred_visited = 0ull;

    // Team-wide minimum of (v, key), the lowest key winning ties; every lane
    // of the team ends with the winner.
    TAMD_HD void team_min(T& v, uint32_t& key) const {
#if defined(__HIP_DEVICE_COMPILE__)
        if constexpr (kT > 1) {
TAMD_UNROLL
            for (int d = kT / 2; d >= 1; d >>= 1) {
                const T ov = __shfl_xor(v, d, kT);
                const uint32_t ok = __shfl_xor(key, d, kT);
                if (ov < v || (ov == v && ok < key)) {
                    v = ov;
                    key = ok;
                }
            }
        }
#else
        (void)v;
        (void)key;
#endif
    }

    // kappa^(-4/5).  On the device in single precision from the hardware
    // base-2 logarithm and exponential (a few ulp; the step size only needs
    // to be a smooth function of the state), elsewhere std::pow.
    TAMD_HD static T kappa_pow(T kappa) {
#if defined(__HIP_DEVICE_COMPILE__)
        if constexpr (std::is_same<T, float>::value) {
            return __builtin_amdgcn_exp2f(-0.8f * __builtin_amdgcn_logf(kappa));
        }
#endif
        return std::pow(kappa, static_cast<T>(-0.8));
    }

    TAMD_HD T step_limit() const {
        const vec3<T> B = field_at(field, st.pos);
        const vec3<T> tb = cross(st.dir, B);
        const T kappa = std::fabs(st.qop) * std::sqrt(dot(tb, tb));
        T h = cfg.step_constraint;
        if (kappa > static_cast<T>(1e-12)) {
            const T ha = cfg.rk_step_scale * kappa_pow(kappa);
            h = ha < h ? ha : h;
        }
        return h;
    }

    TAMD_HD bool do_step(T h) {
        rk4_step<T, kJac>(st, J, field, h, cfg.mass, cfg.abs_q);
        path += h;
        ++n_steps;
        return n_steps <= cfg.max_steps && std::fabs(path) <= cfg.path_limit;
    }

    // Candidate modules of layer `li` in window slots (window position, or
    // module of a plane layer): this lane's slots k = j * kT + tl go to
    // cs[j] / cm[j]; unused slots hold +inf.  No register array is indexed
    // with a runtime value.  s_known: the layer's entry distance from the
    // current state when the caller has it, negative otherwise.
    TAMD_HD void collect(uint32_t li, int exclude, T s_known, T* cs, uint32_t* cm) const {
        const ckf_layer& L = layers[li];
        T s_entry = s_known >= T(0) ? s_known
                                     : layer_entry_distance(L, st.pos, st.dir, cfg.mask_tol_max);
        if (!(s_entry < static_cast<T>(1e29))) {
            s_entry = T(0);
        }
        const vec3<T> q = st.pos + s_entry * st.dir;
        int ic = 0, jc = 0;
        const int nphi = static_cast<int>(L.n_phi);
        const int nlong = static_cast<int>(L.n_long);
        const bool plane = (L.kind == CKF_LAYER_PLANE);
        int wp = cfg.win_phi;
        const int wl = (L.kind == CKF_LAYER_BARREL) ? cfg.win_long : 0;
        if (!plane) {
            layer_bin(L, q, ic, jc);
        }
        if (2 * wp + 1 > nphi) {
            wp = (nphi - 1) / 2;
        }
        const T inf = static_cast<T>(1e30);
        // window slots -> module indices (no memory access yet)
        uint32_t mk[kSlotsPerLane];
        bool okk[kSlotsPerLane];
TAMD_UNROLL
        for (int j = 0; j < kSlotsPerLane; ++j) {
            const int k = j * kT + static_cast<int>(tl);
            bool ok = k < kMaxCandidates;
            uint32_t m = 0u;
            if (plane) {
                ok = ok && k < nphi * nlong;
                m = L.first_surface + static_cast<uint32_t>(k);
            } else {
                const int dp = k / 3 - 1;
                const int dl = k % 3 - 1;
                ok = ok && (dp >= -wp && dp <= wp && dl >= -wl && dl <= wl);
                int ip = (ic + dp) % nphi;
                if (ip < 0) {
                    ip += nphi;
                }
                const int il = jc + dl;
                ok = ok && il >= 0 && il < nlong;
                m = L.first_surface + static_cast<uint32_t>(ip * nlong + il);
            }
            okk[j] = ok && static_cast<int>(m) != exclude;
            mk[j] = okk[j] ? m : L.first_surface;  // a valid module for the loads below
            cm[j] = 0u;
        }
        // straight-line distance to every slot's plane: the centre and normal
        // loads of all slots are independent and go out together
TAMD_UNROLL
        for (int j = 0; j < kSlotsPerLane; ++j) {
            const ckf_surface& S = sfs[mk[j]];
            const vec3<T> w = sf_w<T>(S);
            const T tw = dot(st.dir, w);
            const T s = dot(sf_center<T>(S) - st.pos, w) / tw;
            cs[j] = (okk[j] && std::fabs(tw) >= static_cast<T>(1e-6) && s >= cfg.overstep_tol) ? s : inf;
        }
        // mask test (with the distance-scaled tolerance) for the slots left
TAMD_UNROLL
        for (int j = 0; j < kSlotsPerLane; ++j) {
            if (cs[j] < inf) {
                const ckf_surface& S = sfs[mk[j]];
                const T s = cs[j];
                const vec3<T> x = st.pos + s * st.dir - sf_center<T>(S);
                T tol = cfg.mask_tol_min + cfg.mask_tol_scalor * std::fabs(s);
                tol = tol < cfg.mask_tol_max ? tol : cfg.mask_tol_max;
                if (inside_bounds<T>(S, dot(x, sf_u<T>(S)), dot(x, sf_v<T>(S)), tol)) {
                    cm[j] = mk[j];
                } else {
                    cs[j] = inf;
                }
            }
        }
    }

    // One pass of the step onto the plane of `S`: the signed straight-line
    // distance to the plane decides.  2: plane parallel to the track; 0: within
    // on_surface_tol, h = that distance (the last, sub-tolerance step, none
    // when it is 0); 1: further away, h = the next RK step towards it.
    TAMD_HD int plane_step(const ckf_surface& S, T& h) const {
        const vec3<T> w = sf_w<T>(S);
        const T tw = dot(st.dir, w);
        if (std::fabs(tw) < static_cast<T>(1e-8)) {
            return 2;
        }
        const T sd = -dot(rel_to(st, sf_center<T>(S)), w) / tw;
        if (std::fabs(sd) <= cfg.on_surface_tol) {
            h = sd;
            return 0;
        }
        const T hl = step_limit();
        h = sd > T(0) ? (sd < hl ? sd : hl) : (sd > -hl ? sd : -hl);
        return 1;
    }

    // Step onto the plane of `S`: RK steps until |distance| <= on_surface_tol,
    // then one last (sub-tolerance) RK step of that distance, which keeps the
    // direction, time and transport consistent with the position.
    // 0: on the plane, 1: abort, 2: plane parallel to the track.
    TAMD_HD int step_to_plane(const ckf_surface& S) {
        for (;;) {
            T h = T(0);
            const int r = plane_step(S, h);
            if (r == 2) {
                return 2;
            }
            if (r == 0) {
                return (h != T(0) && !do_step(h)) ? 1 : 0;
            }
            if (!do_step(h)) {
                return 1;
            }
        }
    }

    // The nearest candidate not yet tried (lower module index on ties), over
    // the team; marks it tried.  false when none is left.
    TAMD_HD bool next_candidate(const T* cs, const uint32_t* cm, uint32_t& used,
                                uint32_t& bm) const {
        const T inf = static_cast<T>(1e30);
        T best = inf;
        bm = 0xffffffffu;
TAMD_UNROLL
        for (int j = 0; j < kSlotsPerLane; ++j) {
            const bool avail = ((used >> j) & 1u) == 0u && cs[j] < inf;
            if (avail && (cs[j] < best || (cs[j] == best && cm[j] < bm))) {
                best = cs[j];
                bm = cm[j];
            }
        }
        team_min(best, bm);
        if (!(best < inf)) {
            return false;
        }
TAMD_UNROLL
        for (int j = 0; j < kSlotsPerLane; ++j) {
            if (cs[j] < inf && cm[j] == bm) {
                used |= (1u << j);
            }
        }
        return true;
    }

    // Whether the straight line from the current state still meets the plane
    // of `S` ahead (at least overstep_tol in front).
    TAMD_HD bool plane_ahead(const ckf_surface& S) const {
        const vec3<T> w = sf_w<T>(S);
        const T tw = dot(st.dir, w);
        return !(std::fabs(tw) < static_cast<T>(1e-8) ||
                 dot(sf_center<T>(S) - st.pos, w) / tw < cfg.overstep_tol);
    }

    // On the plane of `S`: inside its mask (within accept_tol), past the
    // minimum path.
    TAMD_HD bool accept(const ckf_surface& S) const {
        const vec3<T> x = rel_to(st, sf_center<T>(S));
        return inside_bounds<T>(S, dot(x, sf_u<T>(S)), dot(x, sf_v<T>(S)), cfg.accept_tol) &&
               path > cfg.min_step_length;
    }

    // The unvisited layer whose (r, z) envelope the straight line from the
    // current state enters first (lowest index on ties, over the team):
    // best = -1 when none; bs = its entry distance.
    TAMD_HD void pick_layer(uint64_t visited, int& best, T& bs) const {
        best = -1;
        bs = static_cast<T>(1e30);
        const T inv_dz = T(1) / st.dir.z;
        if constexpr (kT == 1) {
            if (groups != nullptr && layer_boxes != nullptr && group_boxes != nullptr) {
                pick_layer_grouped(visited, inv_dz, best, bs);
                return;
            }
        }
        for (uint32_t li = tl; li < n_layers; li += kT) {
            if (visited & (1ull << li)) {
                continue;
            }
            // a layer whose entry cannot come before the best one so far
            // is not evaluated (10 um of slack covers the bound's rounding)
            if (layer_entry_lower_bound(layers[li], st.pos, st.dir, cfg.mask_tol_max, inv_dz) >
                bs + static_cast<T>(1e-2)) {
                continue;
            }
            const T s = layer_entry_distance(layers[li], st.pos, st.dir, cfg.mask_tol_max);
            if (s < bs) {
                bs = s;
                best = static_cast<int>(li);
            }
        }
        uint32_t key = best < 0 ? 0xffffffffu : static_cast<uint32_t>(best);
        team_min(bs, key);
        best = key == 0xffffffffu ? -1 : static_cast<int>(key);
    }

    // Entry distance of layer li from the current state (layer_entry_distance;
    // from the layer's box when there is one, the same arithmetic).
    TAMD_HD T entry_distance(uint32_t li) const {
        if (layer_boxes != nullptr) {
            const nav_box B = layer_boxes[li];
            if (B.kind == kBoxEnvelope) {
                return envelope_entry_distance<T>(static_cast<T>(B.r0), static_cast<T>(B.r1),
                                                  static_cast<T>(B.z0), static_cast<T>(B.z1),
                                                  st.pos, st.dir, cfg.mask_tol_max);
            }
        }
        return layer_entry_distance(layers[li], st.pos, st.dir, cfg.mask_tol_max);
    }

    // pick_layer over layer groups, best first: the bound of every group (four
    // boxes loaded at a time, no branches), then rounds that take the group
    // with the smallest bound left and evaluate its members -- each with its own
    // box bound and, when that does not rule it out, its entry distance; the
    // search ends when the smallest bound left is past the best distance.  A
    // group or layer is only passed over when its bound exceeds the best
    // distance (with the scan's 10 um of slack) or it cannot be reached, and
    // equal distances go to the lower layer index, so the result is the
    // scan's.
    TAMD_HD void pick_layer_grouped(uint64_t visited, T inv_dz, int& best, T& bs) const {
        const T big = static_cast<T>(1e29);
        const T slack = static_cast<T>(1e-2);
        const T mt = cfg.mask_tol_max;
        (void)inv_dz;
        const nav_ray<T> q = make_nav_ray(st.pos, st.dir);
        // every group's bound, once (four boxes loaded at a time)
        T gb[kMaxGroups];
TAMD_UNROLL
        for (uint32_t g0 = 0; g0 < kMaxGroups; g0 += 4u) {
            nav_box bx[4];
TAMD_UNROLL
            for (uint32_t j = 0; j < 4u; ++j) {
                bx[j] = group_boxes[g0 + j];  // slots past n_groups: kBoxEmpty
            }
TAMD_UNROLL
            for (uint32_t j = 0; j < 4u; ++j) {
                gb[g0 + j] = nav_box_bound<T>(bx[j], q, mt);
            }
        }
        for (uint32_t round = 0; round < n_groups; ++round) {
            // the smallest and second smallest bound of the groups left
            T m1 = big, m2 = big;
            uint32_t g1 = kMaxGroups;
TAMD_UNROLL
            for (uint32_t g = 0; g < kMaxGroups; ++g) {
                const T b = gb[g];
                const bool lt1 = b < m1;
                const bool lt2 = b < m2;
                m2 = lt1 ? m1 : (lt2 ? b : m2);
                g1 = lt1 ? g : g1;
                m1 = lt1 ? b : m1;
            }
            if (g1 == kMaxGroups || m1 > bs + slack) {
                return;
            }
TAMD_UNROLL
            for (uint32_t g = 0; g < kMaxGroups; ++g) {
                gb[g] = g == g1 ? static_cast<T>(1e30) : gb[g];
            }
            const layer_group& G = groups[g1];
            const uint32_t gn = G.n;
            uint32_t mem[kGroupSize];
            nav_box mb[kGroupSize];
TAMD_UNROLL
            for (uint32_t k = 0; k < kGroupSize; ++k) {
                mem[k] = G.member[k < gn ? k : 0u];
                mb[k] = layer_boxes[mem[k]];
            }
            uint32_t cand = 0u;
TAMD_UNROLL
            for (uint32_t k = 0; k < kGroupSize; ++k) {
                const T lb = nav_box_bound<T>(mb[k], q, mt);
                const bool ok = k < gn && ((visited >> mem[k]) & 1ull) == 0ull && lb < big &&
                                !(lb > bs + slack);
                cand |= ok ? (1u << k) : 0u;
            }
            while (cand != 0u) {
                const uint32_t k = static_cast<uint32_t>(__builtin_ctz(cand));
                cand &= cand - 1u;
                const uint32_t li = G.member[k];
                const T s = entry_distance(li);
                if (s < bs || (s == bs && static_cast<int>(li) < best)) {
                    bs = s;
                    best = static_cast<int>(li);
                }
            }
            // every group left is bounded by m2 or more
            if (!(m2 <= bs + slack)) {
                return;
            }
        }
    }

    // Below min_p or min_pT at the start: the propagation aborts at once.
    TAMD_HD bool below_momentum_cut() const {
        const T p = cfg.abs_q / std::fabs(st.qop);
        const T pT = p * std::sqrt(st.dir.x * st.dir.x + st.dir.y * st.dir.y);
        return p < cfg.min_p || pT < cfg.min_pT;
    }

    // Visit the candidates of layer `li` in increasing (distance, module).
    // 0: found (out_sf), 1: nothing in this layer, 2: abort.
    TAMD_HD int process_layer(uint32_t li, int exclude, T s_known, uint32_t& out_sf) {
        T cs[kSlotsPerLane];
        uint32_t cm[kSlotsPerLane];
        collect(li, exclude, s_known, cs, cm);
        uint32_t used = 0u;
        for (int tries = 0; tries < kMaxCandidates; ++tries) {
            uint32_t bm = 0u;
            if (!next_candidate(cs, cm, used, bm)) {
                return 1;
            }
            const ckf_surface& S = sfs[bm];
            if (!plane_ahead(S)) {
                continue;
            }
            const int r = step_to_plane(S);
            if (r == 1) {
                return 2;
            }
            if (r == 2) {
                continue;
            }
            if (accept(S)) {
                out_sf = bm;
                return 0;
            }
        }
        return 1;
    }

    // Propagate from the current state (on surface start_sf, or -1 for a
    // free start) to the next sensitive surface.  With defer, a propagation
    // that needs a second layer search stops there with NAV_DEFER (state,
    // path and step count kept, the visited layers in deferred_visited).
    TAMD_HD int propagate(int start_sf, uint32_t& out_sf, bool defer = false) {
        path = T(0);
        n_steps = 0u;
        if (below_momentum_cut()) {
            return NAV_ABORT;
        }
        uint64_t visited = 0ull;
        if (start_sf >= 0) {
            const uint32_t l0 = sfs[start_sf].layer;
            const int r = process_layer(l0, start_sf, T(-1), out_sf);
            if (r == 0) {
                return NAV_FOUND;
            }
            if (r == 2) {
                return NAV_ABORT;
            }
            visited |= (1ull << l0);
        }
        return layer_loop(visited, 0u, out_sf, defer);
    }

    // The layer loop of propagate from iteration guard0 with the layers in
    // `visited` done: pick the next layer, approach it, visit its window.
    TAMD_HD int layer_loop(uint64_t visited, uint32_t guard0, uint32_t& out_sf, bool defer) {
        for (uint32_t guard = guard0; guard <= n_layers; ++guard) {
            if (defer && guard > guard0) {
                deferred_visited = visited;
                return NAV_DEFER;
            }
            int best = -1;
            T bs = T(0);
            pick_layer(visited, best, bs);
            if (best < 0 || !(bs < static_cast<T>(1e29))) {
                return NAV_EXIT;
            }
            while (bs > cfg.approach_dist && bs < static_cast<T>(1e29)) {
                const T hl = step_limit();
                if (!do_step(bs < hl ? bs : hl)) {
                    return NAV_ABORT;
                }
                bs = entry_distance(static_cast<uint32_t>(best));
            }
            if (bs < static_cast<T>(1e29)) {
                // bs is the entry distance from the current state (recomputed
                // after every approach step)
                const int r = process_layer(static_cast<uint32_t>(best), -1, bs, out_sf);
                if (r == 0) {
                    return NAV_FOUND;
                }
                if (r == 2) {
                    return NAV_ABORT;
                }
            }
            visited |= (1ull << best);
        }
        return NAV_EXIT;
    }
};

}  // namespace tamd
