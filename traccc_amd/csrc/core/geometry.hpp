// Toy cylindrical detector geometry: planar modules (rectangles on barrel
// layers, trapezoids on endcap rings) grouped in navigation layers, each with
// an (r, z) envelope and a regular (phi, long) module grid.
//
// This stands in for the detray toy detector the reference builds its tests
// on (tests/common/tests/toy_detector_fixture.hpp:58-82); detray itself is not
// available, so the surface/mask/grid layout is this project's own.
#pragma once

#include <cmath>

#include "defs.hpp"

namespace tamd {

template <typename T>
TAMD_HD vec3<T> sf_center(const ckf_surface& s) {
    return {static_cast<T>(s.center[0]), static_cast<T>(s.center[1]),
            static_cast<T>(s.center[2])};
}
template <typename T>
TAMD_HD vec3<T> sf_u(const ckf_surface& s) {
    return {static_cast<T>(s.axis_u[0]), static_cast<T>(s.axis_u[1]),
            static_cast<T>(s.axis_u[2])};
}
template <typename T>
TAMD_HD vec3<T> sf_v(const ckf_surface& s) {
    return {static_cast<T>(s.axis_v[0]), static_cast<T>(s.axis_v[1]),
            static_cast<T>(s.axis_v[2])};
}
template <typename T>
TAMD_HD vec3<T> sf_w(const ckf_surface& s) {
    return {static_cast<T>(s.normal[0]), static_cast<T>(s.normal[1]),
            static_cast<T>(s.normal[2])};
}

// Mask check in local coordinates with a tolerance (0 = exact bounds).
template <typename T>
TAMD_HD bool inside_bounds(const ckf_surface& s, T l0, T l1, T tol) {
    const T b0 = static_cast<T>(s.bound[0]);
    const T b1 = static_cast<T>(s.bound[1]);
    if (s.bound_kind == CKF_BOUND_RECTANGLE) {
        return std::fabs(l0) <= b0 + tol && std::fabs(l1) <= b1 + tol;
    }
    const T h = static_cast<T>(s.bound[2]);
    if (std::fabs(l1) > h + tol) {
        return false;
    }
    const T half = b0 + (b1 - b0) * (l1 + h) / (T(2) * h);
    return std::fabs(l0) <= half + tol;
}

// Straight-line distance along d at which the line enters the layer's (r, z)
// envelope; 0 when p is already inside, +inf when the line never enters.
// Crossing points are accepted within `margin` of the envelope's other extent.
template <typename T>
TAMD_HD T layer_entry_distance(const ckf_layer& L, const vec3<T>& p,
                               const vec3<T>& d, T margin) {
    const T inf = static_cast<T>(1e30);
    if (L.kind == CKF_LAYER_PLANE) {
        const vec3<T> n{static_cast<T>(L.normal[0]), static_cast<T>(L.normal[1]),
                        static_cast<T>(L.normal[2])};
        const T d0 = static_cast<T>(L.env_r[0]);
        const T d1 = static_cast<T>(L.env_r[1]);
        const T dp = dot(n, p);
        if (dp >= d0 && dp <= d1) {
            return T(0);
        }
        const T dn = dot(n, d);
        if (std::fabs(dn) < static_cast<T>(1e-12)) {
            return inf;
        }
        const T s0 = (d0 - dp) / dn;
        const T s1 = (d1 - dp) / dn;
        T best = inf;
        if (s0 > T(0)) {
            best = s0;
        }
        if (s1 > T(0) && s1 < best) {
            best = s1;
        }
        return best;
    }
    const T r0 = static_cast<T>(L.env_r[0]);
    const T r1 = static_cast<T>(L.env_r[1]);
    const T z0 = static_cast<T>(L.env_z[0]);
    const T z1 = static_cast<T>(L.env_z[1]);
    // Envelopes the line cannot reach, before any division: beyond the
    // widened z range and not moving towards it (z(s) only moves away, every
    // wall crossing then fails its z test and both z faces lie behind), or
    // outside the widened outer radius and not moving inwards (r(s) only
    // grows, so every crossing fails its radius test).  Both give inf, as
    // the full computation below would.
    if ((p.z > z1 + margin && d.z >= T(0)) || (p.z < z0 - margin && d.z <= T(0))) {
        return inf;
    }
    const T rr = p.x * p.x + p.y * p.y;
    const T hb0 = p.x * d.x + p.y * d.y;
    if (rr > (r1 + margin) * (r1 + margin) && hb0 >= T(0)) {
        return inf;
    }
    if (rr >= r0 * r0 && rr <= r1 * r1 && p.z >= z0 && p.z <= z1) {
        return T(0);
    }
    T best = inf;
    const T a = d.x * d.x + d.y * d.y;
    if (a > static_cast<T>(1e-12)) {
        const T hb = hb0;
        const T radii[2] = {r0, r1};
        for (int k = 0; k < 2; ++k) {
            const T c = rr - radii[k] * radii[k];
            const T disc = hb * hb - a * c;
            if (disc >= T(0)) {
                const T sq = std::sqrt(disc);
                const T roots[2] = {(-hb - sq) / a, (-hb + sq) / a};
                for (int j = 0; j < 2; ++j) {
                    const T s = roots[j];
                    if (s > T(0) && s < best) {
                        const T z = p.z + s * d.z;
                        if (z >= z0 - margin && z <= z1 + margin) {
                            best = s;
                        }
                    }
                }
            }
        }
    }
    if (std::fabs(d.z) > static_cast<T>(1e-12)) {
        const T zs[2] = {z0, z1};
        for (int k = 0; k < 2; ++k) {
            const T s = (zs[k] - p.z) / d.z;
            if (s > T(0) && s < best) {
                const T x = p.x + s * d.x;
                const T y = p.y + s * d.y;
                const T r2 = x * x + y * y;
                const T lo = r0 - margin > T(0) ? r0 - margin : T(0);
                const T hi = r1 + margin;
                if (r2 >= lo * lo && r2 <= hi * hi) {
                    best = s;
                }
            }
        }
    }
    return best;
}

// A lower bound on layer_entry_distance(L, p, d, margin) for an (r, z)
// envelope, from the slab the line must reach first: an accepted crossing
// lies within the widened z range and at a radius of at least r0 - margin,
// and the straight line covers at most its path length in z and in radius.
// inv_dz is 1 / d.z (inf for d.z == 0).  0 when nothing is known.
template <typename T>
TAMD_HD T layer_entry_lower_bound(const ckf_layer& L, const vec3<T>& p, const vec3<T>& d,
                                  T margin, T inv_dz) {
    if (L.kind == CKF_LAYER_PLANE) {
        return T(0);
    }
    T lb = T(0);
    const T zlo = static_cast<T>(L.env_z[0]) - margin;
    const T zhi = static_cast<T>(L.env_z[1]) + margin;
    if (p.z < zlo && d.z > T(0)) {
        lb = (zlo - p.z) * inv_dz;
    } else if (p.z > zhi && d.z < T(0)) {
        lb = (zhi - p.z) * inv_dz;
    }
    const T rlo = static_cast<T>(L.env_r[0]) - margin;
    const T rr = p.x * p.x + p.y * p.y;
    if (rlo > T(0) && rr < rlo * rlo) {
        // r(s) <= r(0) + s and r(0) < rlo: s >= rlo - r(0) >= (rlo^2 - r(0)^2) / (2 rlo)
        const T lr = (rlo * rlo - rr) / (T(2) * rlo);
        lb = lr > lb ? lr : lb;
    }
    return lb;
}

// Module grid bin of a global point: phi bin (wrapped) and long bin (clamped).
template <typename T>
TAMD_HD void layer_bin(const ckf_layer& L, const vec3<T>& q, int& iphi,
                       int& il) {
    const T phi = std::atan2(q.y, q.x);
    const T dphi = static_cast<T>(kTwoPi) / static_cast<T>(L.n_phi);
    const T f = (phi - static_cast<T>(L.phi0)) / dphi;
    int ic = static_cast<int>(std::floor(f + T(0.5)));
    const int n = static_cast<int>(L.n_phi);
    ic %= n;
    if (ic < 0) {
        ic += n;
    }
    iphi = ic;
    if (L.kind == CKF_LAYER_BARREL && L.n_long > 1u) {
        const T g = (q.z - static_cast<T>(L.long0)) /
                    static_cast<T>(L.long_pitch);
        int jl = static_cast<int>(std::floor(g + T(0.5)));
        const int nl = static_cast<int>(L.n_long);
        jl = jl < 0 ? 0 : (jl >= nl ? nl - 1 : jl);
        il = jl;
    } else {
        il = 0;
    }
}

}  // namespace tamd
