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

// Straight-line distance along d at which the line enters the (r, z)
// envelope [r0, r1] x [z0, z1]; 0 when p is already inside, +inf when the line
// never enters.  Crossing points are accepted within `margin` of the
// envelope's other extent.
template <typename T>
TAMD_HD T envelope_entry_distance(T r0, T r1, T z0, T z1, const vec3<T>& p, const vec3<T>& d,
                                  T margin) {
    const T inf = static_cast<T>(1e30);
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

// Straight-line distance along d at which the line enters the layer's
// envelope (envelope_entry_distance; a plane layer: its slab).
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
    return envelope_entry_distance<T>(static_cast<T>(L.env_r[0]), static_cast<T>(L.env_r[1]),
                                      static_cast<T>(L.env_z[0]), static_cast<T>(L.env_z[1]), p,
                                      d, margin);
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

// Navigation layer groups: up to four layers under one (r, z) box that
// contains all their envelopes, so that a single bound can rule out every
// member.  Plane layers go in groups without a box (unbounded).
constexpr uint32_t kGroupSize = 4u;
constexpr uint32_t kMaxGroups = 16u;  // 64 layers

struct layer_group {
    float env_r[2];  // union of the members' (r, z) envelopes
    float env_z[2];
    uint32_t bounded;  // 0: plane layers, no box
    uint32_t n;
    uint32_t member[kGroupSize];
};

// A layer's or a group's (r, z) box with the margin-dependent terms of its
// lower bound, for one navigation configuration: 32 B, two 16-byte loads.
enum nav_box_kind : uint32_t { kBoxEnvelope = 0u, kBoxNone = 1u, kBoxEmpty = 2u };

struct alignas(16) nav_box {
    float r0, r1, z0, z1;  // the envelope (a group: the union of its members')
    float r1m2;            // (r1 + margin)^2
    uint32_t kind;         // kBoxEnvelope; kBoxNone (plane layers: bound 0); kBoxEmpty
    float pad[2];
};

TAMD_HD nav_box make_nav_box(float r0, float r1, float z0, float z1, uint32_t kind, float margin) {
    nav_box B;
    B.r0 = r0;
    B.r1 = r1;
    B.z0 = z0;
    B.z1 = z1;
    B.r1m2 = (r1 + margin) * (r1 + margin);
    B.kind = kind;
    B.pad[0] = 0.f;
    B.pad[1] = 0.f;
    return B;
}

// A straight line from p along the unit vector d with the terms every box
// bound shares: rr = |p_T|^2, hb0 = p_T . d_T, r = |p_T|, inv_dt = 1 / |d_T|
// and inv_dz = 1 / d.z (inf when zero).
template <typename T>
struct nav_ray {
    vec3<T> p, d;
    T rr, hb0, r, inv_dt, inv_dz;
};

template <typename T>
TAMD_HD nav_ray<T> make_nav_ray(const vec3<T>& p, const vec3<T>& d) {
    nav_ray<T> q;
    q.p = p;
    q.d = d;
    q.rr = p.x * p.x + p.y * p.y;
    q.hb0 = p.x * d.x + p.y * d.y;
    q.r = std::sqrt(q.rr);
    q.inv_dt = T(1) / std::sqrt(d.x * d.x + d.y * d.y);
    q.inv_dz = T(1) / d.z;
    return q;
}

// A lower bound on envelope_entry_distance over every envelope inside box B,
// without branches.  An accepted crossing lies in the box widened by the
// margin: z within [z0 - m, z1 + m] and radius at least r0 - m.  The bound is
// the larger of the z-slab bound (the z distance to the widened range over
// |d.z|) and the radial bound (the radius grows by at most |d_T| per unit
// path, so reaching r0 - m from r takes at least (r0 - m - r) / |d_T|); it
// is 1e30 when the line cannot reach the widened box: beyond its z range and
// not moving towards it, or outside its outer radius (with a relative safety
// margin) and not moving inwards -- the members' own distances are +inf by
// the same comparisons.  A larger box gives a smaller bound.  kBoxNone: 0;
// kBoxEmpty: 1e30.
template <typename T>
TAMD_HD T nav_box_bound(const nav_box& B, const nav_ray<T>& q, T margin) {
    const T zlo = static_cast<T>(B.z0) - margin;
    const T zhi = static_cast<T>(B.z1) + margin;
    const bool below = q.p.z < zlo;
    const bool above = q.p.z > zhi;
    const bool gone = (above && q.d.z >= T(0)) || (below && q.d.z <= T(0)) ||
                      (q.rr > static_cast<T>(B.r1m2) * static_cast<T>(1.000001f) && q.hb0 >= T(0));
    const T lz = below ? (zlo - q.p.z) * q.inv_dz : (above ? (zhi - q.p.z) * q.inv_dz : T(0));
    const T rlo = static_cast<T>(B.r0) - margin;
    const T lr = rlo > q.r ? (rlo - q.r) * q.inv_dt : T(0);
    T lb = lz > lr ? lz : lr;
    lb = gone ? static_cast<T>(1e30) : lb;
    lb = B.kind == kBoxNone ? T(0) : lb;
    return B.kind == kBoxEmpty ? static_cast<T>(1e30) : lb;
}

// The boxes of the layers ([n_layers]) and of the groups (kMaxGroups slots,
// kBoxEmpty past n_groups) for a margin; one record per index k < n_layers +
// kMaxGroups (layers first), so a block of threads can fill them together.
TAMD_HD nav_box nav_box_of(const ckf_layer* layers, uint32_t n_layers, const layer_group* groups,
                           uint32_t n_groups, float margin, uint32_t k) {
    if (k < n_layers) {
        const ckf_layer& L = layers[k];
        return make_nav_box(L.env_r[0], L.env_r[1], L.env_z[0], L.env_z[1],
                            L.kind == CKF_LAYER_PLANE ? kBoxNone : kBoxEnvelope, margin);
    }
    const uint32_t g = k - n_layers;
    if (g >= n_groups) {
        return make_nav_box(0.f, 0.f, 0.f, 0.f, kBoxEmpty, margin);
    }
    const layer_group& G = groups[g];
    return make_nav_box(G.env_r[0], G.env_r[1], G.env_z[0], G.env_z[1],
                        G.bounded != 0u ? kBoxEnvelope : kBoxNone, margin);
}

// Host: group the layers for the navigator.  (r, z) envelope layers are
// ordered by (kind, z centre, r centre); a group takes consecutive layers of
// one kind while the next one's z range lies within kGroupGap of the
// group's (a barrel's radii, a disk's rings) and it holds fewer than
// kGroupSize.  Plane layers follow in runs of their own.  Should that give
// more than kMaxGroups groups, the ordered layers are cut into plain runs of
// kGroupSize instead (at most 16 for 64 layers).  Returns the number of
// groups.
constexpr float kGroupGap = 20.f;  // mm

inline uint32_t build_layer_groups(const ckf_layer* L, uint32_t n, layer_group* out) {
    uint32_t order[64];
    uint32_t m = 0u;
    for (uint32_t i = 0; i < n && i < 64u; ++i) {
        if (L[i].kind != CKF_LAYER_PLANE) order[m++] = i;
    }
    const uint32_t n_box = m;
    for (uint32_t i = 0; i < n && i < 64u; ++i) {
        if (L[i].kind == CKF_LAYER_PLANE) order[m++] = i;
    }
    // insertion sort of the envelope layers (at most 64)
    auto key_less = [&](uint32_t a, uint32_t b) {
        if (L[a].kind != L[b].kind) return L[a].kind < L[b].kind;
        const float za = L[a].env_z[0] + L[a].env_z[1], zb = L[b].env_z[0] + L[b].env_z[1];
        if (za != zb) return za < zb;
        const float ra = L[a].env_r[0] + L[a].env_r[1], rb = L[b].env_r[0] + L[b].env_r[1];
        if (ra != rb) return ra < rb;
        return a < b;
    };
    for (uint32_t i = 1; i < n_box; ++i) {
        const uint32_t v = order[i];
        uint32_t j = i;
        while (j > 0u && key_less(v, order[j - 1u])) {
            order[j] = order[j - 1u];
            --j;
        }
        order[j] = v;
    }
    for (int pass = 0; pass < 2; ++pass) {
        const bool clustered = pass == 0;
        uint32_t g = 0u;
        for (uint32_t start = 0; start < m;) {
            // a group never mixes envelope and plane layers
            const uint32_t end_run = start < n_box ? n_box : m;
            layer_group G{};
            G.bounded = start < n_box ? 1u : 0u;
            G.env_r[0] = 1e30f;
            G.env_r[1] = -1e30f;
            G.env_z[0] = 1e30f;
            G.env_z[1] = -1e30f;
            uint32_t k = start;
            while (k < end_run && G.n < kGroupSize) {
                const ckf_layer& l = L[order[k]];
                if (clustered && G.n > 0u && G.bounded != 0u &&
                    (l.kind != L[order[start]].kind || l.env_z[0] > G.env_z[1] + kGroupGap ||
                     l.env_z[1] < G.env_z[0] - kGroupGap)) {
                    break;
                }
                G.member[G.n++] = order[k];
                G.env_r[0] = l.env_r[0] < G.env_r[0] ? l.env_r[0] : G.env_r[0];
                G.env_r[1] = l.env_r[1] > G.env_r[1] ? l.env_r[1] : G.env_r[1];
                G.env_z[0] = l.env_z[0] < G.env_z[0] ? l.env_z[0] : G.env_z[0];
                G.env_z[1] = l.env_z[1] > G.env_z[1] ? l.env_z[1] : G.env_z[1];
                ++k;
            }
            if (g < kMaxGroups) {
                out[g] = G;
            }
            ++g;
            start = k;
        }
        if (g <= kMaxGroups) {
            return g;
        }
    }
    return kMaxGroups;  // not reached: plain runs of four give at most 16 groups
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
