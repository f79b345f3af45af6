// Bound-state I/O and transport shared by the CKF step kernels and the
// progressive filter: structure-of-arrays state columns, seeds, the
// pointwise multiple-scattering noise of a module, and the transport of a
// bound state with its covariance to the next sensitive surface.
#pragma once

#include <hip/hip_runtime.h>

#include "../core/propagator.hpp"
#include "device_types.hpp"

namespace tamd {

__device__ inline void load_soa(const float* __restrict__ base, uint32_t cap, uint32_t i, float* x,
                                float* c, uint32_t& sf) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        x[k] = base[static_cast<size_t>(k) * cap + i];
    }
#pragma unroll
    for (int k = 0; k < 21; ++k) {
        c[k] = base[static_cast<size_t>(6 + k) * cap + i];
    }
    sf = __float_as_uint(base[static_cast<size_t>(27) * cap + i]);
}

__device__ inline void store_soa(float* __restrict__ base, uint32_t cap, uint32_t i, const float* x,
                                 const float* c, uint32_t sf) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        base[static_cast<size_t>(k) * cap + i] = x[k];
    }
#pragma unroll
    for (int k = 0; k < 21; ++k) {
        base[static_cast<size_t>(6 + k) * cap + i] = c[k];
    }
    base[static_cast<size_t>(27) * cap + i] = __uint_as_float(sf);
}

__device__ inline void load_seed(const ckf_params& P, uint32_t s, float* x, float* c, uint32_t& sf) {
    const float* v = P.seed_vec + 6u * static_cast<size_t>(s);
    const float* m = P.seed_cov + 36u * static_cast<size_t>(s);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        x[k] = v[k];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = i; j < 6; ++j) {
            c[sym_idx(i, j)] = m[6 * i + j];
        }
    }
    sf = P.seed_surface[s];
}

__device__ inline void load_ref(const ckf_params& P, uint32_t ref, uint32_t in_parity, float* x,
                                float* c, uint32_t& sf) {
    const uint32_t kind = ref >> kRefKindShift;
    const uint32_t idx = ref & kRefMask;
    if (kind == kRefFiltered) {
        load_soa(P.flt, P.flt_cap, idx, x, c, sf);
    } else if (kind == kRefInput) {
        load_soa(P.state[in_parity], P.state_cap, idx, x, c, sf);
    } else {
        load_seed(P, idx, x, c, sf);
    }
}

// Multiple-scattering noise of module S for a state on it (covariance part
// of the pointwise material interaction).
__device__ inline void add_material_sf(const ckf_surface& S, float mass, float abs_q,
                                       const float* x, float* c) {
    const float t = S.thickness_x0;
    if (!(t > 0.f)) {
        return;
    }
    float sp, cp, st, ct;
    sincosf(x[e_phi], &sp, &cp);
    sincosf(x[e_theta], &st, &ct);
    const float cosi = fabsf(cp * st * S.normal[0] + sp * st * S.normal[1] + ct * S.normal[2]);
    const float t0 = highland_theta0<float>(t / cosi, x[e_qop], mass, abs_q);
    const float v = t0 * t0;
    c[sym_idx(e_phi, e_phi)] += v / (st * st);
    c[sym_idx(e_theta, e_theta)] += v;
}

__device__ inline void add_material(const ckf_params& P, uint32_t sf, const float* x, float* c) {
    add_material_sf(P.sfs[sf], P.nav.mass, P.nav.abs_q, x, c);
}

// The navigation layers, their groups and the boxes of both (for the
// configuration's mask tolerance), staged in LDS by every block of a
// transport kernel (64 layers and 16 groups at most).
struct nav_tables {
    nav_box boxes[64 + kMaxGroups];  // layers [0, n_layers), then the group slots
    ckf_layer layers[64];
    layer_group groups[kMaxGroups];
    uint32_t n_layers;
};

__device__ inline void stage_nav_tables(nav_tables& t, const ckf_layer* layers, uint32_t n_layers,
                                        const layer_group* groups, uint32_t n_groups,
                                        float margin) {
    for (uint32_t k = threadIdx.x; k < n_layers; k += blockDim.x) {
        t.layers[k] = layers[k];
    }
    for (uint32_t k = threadIdx.x; k < n_groups; k += blockDim.x) {
        t.groups[k] = groups[k];
    }
    for (uint32_t k = threadIdx.x; k < n_layers + kMaxGroups; k += blockDim.x) {
        t.boxes[k < n_layers ? k : 64u + (k - n_layers)] =
            nav_box_of(layers, n_layers, groups, n_groups, margin, k);
    }
    if (threadIdx.x == 0u) {
        t.n_layers = n_layers;
    }
    __syncthreads();
}

// Geometry and field a transport needs.
struct geo_view {
    const ckf_surface* sfs;
    uint32_t n_layers;
    field_view<float> field;
};

// c <- Jb c Jb^T on the packed symmetric covariance.
__device__ inline void transport_cov(const float (&Jb)[6][6], float* c) {
    float M[6][6];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc += Jb[r][k] * c[sym_idx(k, q)];
            M[r][q] = acc;
        }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
#pragma unroll
        for (int q = r; q < 6; ++q) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc += M[r][k] * Jb[q][k];
            c[sym_idx(r, q)] = acc;
        }
    }
}

// A navigator over geometry G for one transport (kT lanes per track, tl
// this lane's index in the team), with the LDS navigation tables when given.
template <int kT = 1>
__device__ inline navigator<float, true, kT> make_navigator(const geo_view& G, const ckf_layer* layers,
                                                          const nav_config<float>& cfg, uint32_t tl,
                                                          const nav_tables* tabs,
                                                          uint32_t n_groups) {
    navigator<float, true, kT> nav;
    nav.tl = tl;
    nav.sfs = G.sfs;
    nav.layers = layers;
    nav.n_layers = G.n_layers;
    if (tabs != nullptr) {
        nav.groups = tabs->groups;
        nav.n_groups = n_groups;
        nav.layer_boxes = tabs->boxes;
        nav.group_boxes = tabs->boxes + 64;
    }
    nav.field = G.field;
    nav.cfg = cfg;
    return nav;
}

// The bound state and covariance on the navigator's destination surface
// end_sf after a NAV_FOUND propagation: x receives the bound vector, c (the
// covariance at the start) is transported.
template <int kT>
__device__ inline void finish_transport(const geo_view& G, const nav_config<float>& cfg,
                                        const navigator<float, true, kT>& nav, uint32_t end_sf,
                                        float* x, float* c) {
    float nx[6];
    float Jb[6][6];
    free_to_bound<float, true>(G.sfs[end_sf], nav.st, nav.J, G.field, cfg.mass, cfg.abs_q, nx, Jb);
    transport_cov(Jb, c);
#pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = nx[k];
}

// Transport (x, c) from surface sf to the next sensitive surface
// (propagate_to_next_surface's stepper, navigator and parameter transport,
// without the material of the destination).  On NAV_FOUND x, c and sf hold
// the state on the new surface; otherwise they are unchanged.  n_steps and
// path receive the RK steps and the path length of the leg.  With kT > 1 the
// kT lanes of a team call this together with the same state (tl: the lane's
// index in the team) and all get the same result.
template <int kT = 1>
__device__ inline int transport_bound(const geo_view& G, const ckf_layer* layers,
                                      const nav_config<float>& cfg, float* x, float* c,
                                      uint32_t& sf, uint32_t& n_steps, float& path,
                                      uint32_t tl = 0u, const nav_tables* tabs = nullptr,
                                      uint32_t n_groups = 0u) {
    navigator<float, true, kT> nav = make_navigator<kT>(G, layers, cfg, tl, tabs, n_groups);
    bound_to_free<float, true>(G.sfs[sf], x, nav.st, nav.J);
    uint32_t end_sf = 0u;
    const int status = nav.propagate(static_cast<int>(sf), end_sf);
    n_steps = nav.n_steps;
    path = nav.path;
    if (status != NAV_FOUND) {
        return status;
    }
    finish_transport<kT>(G, cfg, nav, end_sf, x, c);
    sf = end_sf;
    return NAV_FOUND;
}

template <int kT = 1>
__device__ inline int transport_state(const ckf_params& P, const ckf_layer* layers,
                                      const nav_config<float>& cfg, float* x, float* c,
                                      uint32_t& sf, uint32_t& n_steps, float& path,
                                      uint32_t tl = 0u, const nav_tables* tabs = nullptr,
                                      uint32_t n_groups = 0u) {
    const geo_view G{P.sfs, P.n_layers, P.field};
    return transport_bound<kT>(G, layers, cfg, x, c, sf, n_steps, path, tl, tabs, n_groups);
}

// Transport (x, c) from surface sf straight onto surface `target` (forwards
// or backwards along the track, no navigation: the surface sequence of a
// fit).  0 on success (x, c, sf updated), NAV_ABORT when the step or path
// limit is hit or the plane is parallel to the track (x, c, sf unchanged).
__device__ inline int transport_to_surface(const geo_view& G, const nav_config<float>& cfg,
                                           float* x, float* c, uint32_t& sf, uint32_t target) {
    navigator<float, true, 1> nav;
    nav.tl = 0u;
    nav.sfs = G.sfs;
    nav.layers = nullptr;
    nav.n_layers = G.n_layers;
    nav.field = G.field;
    nav.cfg = cfg;
    nav.path = 0.f;
    nav.n_steps = 0u;
    bound_to_free<float, true>(G.sfs[sf], x, nav.st, nav.J);
    if (nav.step_to_plane(G.sfs[target]) != 0) {
        return NAV_ABORT;
    }
    float nx[6];
    float Jb[6][6];
    free_to_bound<float, true>(G.sfs[target], nav.st, nav.J, G.field, cfg.mass, cfg.abs_q, nx, Jb);
    transport_cov(Jb, c);
#pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = nx[k];
    sf = target;
    return 0;
}

}  // namespace tamd
