// Synthetic-input toolkit: toy detectors, the solenoid field map and seeded
// particle-gun events (hits from the same navigator the CKF uses, Highland
// scattering, Gaussian smearing, noise and dead modules).
//
// Behavioural references: traccc's fast simulator and smearer
// (simulation/include/traccc/simulation/simulator.hpp:99-153,
//  simulation/include/traccc/simulation/measurement_smearer.hpp:26-124),
// truth seeds on the first measurement's surface
// (performance/include/traccc/utils/event_data.hpp:79-118,
//  core/include/traccc/utils/seed_generator.hpp:87-184), and the measurement
// ordering (core/include/traccc/edm/impl/measurement_collection.ipp:41-57).
#include "traccc_amd/simulation.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "../core/propagator.hpp"

namespace {

using tamd::kPi;
using tamd::kTwoPi;

// ---------------------------------------------------------------------------
// Random numbers: xoshiro256** seeded through splitmix64 (Blackman & Vigna),
// Box-Muller normals, Knuth Poisson.
// ---------------------------------------------------------------------------
struct rng {
    uint64_t s[4];
    explicit rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) {
            x += 0x9e3779b97f4a7c15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // (0, 1)
    double uniform() { return (static_cast<double>(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        const double u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
    }
    uint32_t poisson(double lambda) {
        if (lambda <= 0.0) {
            return 0u;
        }
        const double l = std::exp(-lambda);
        uint32_t k = 0;
        double p = 1.0;
        do {
            ++k;
            p *= uniform();
        } while (p > l);
        return k - 1u;
    }
};

// ---------------------------------------------------------------------------
// Detector presets
// ---------------------------------------------------------------------------
struct det_builder {
    std::vector<ckf_surface> sfs;
    std::vector<ckf_layer> layers;

    void set_frame(ckf_surface& s, double cx, double cy, double cz, const double u[3],
                   const double v[3]) {
        s.center[0] = static_cast<float>(cx);
        s.center[1] = static_cast<float>(cy);
        s.center[2] = static_cast<float>(cz);
        const double w[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                             u[0] * v[1] - u[1] * v[0]};
        for (int i = 0; i < 3; ++i) {
            s.axis_u[i] = static_cast<float>(u[i]);
            s.axis_v[i] = static_cast<float>(v[i]);
            s.normal[i] = static_cast<float>(w[i]);
        }
    }

    ckf_surface& new_surface(uint32_t layer, uint32_t local_index) {
        sfs.emplace_back();
        ckf_surface& s = sfs.back();
        std::memset(&s, 0, sizeof(s));
        s.layer = layer;
        s.sensitive = 1u;
        s.identifier = (static_cast<uint64_t>(layer + 1u) << 24) | local_index;
        return s;
    }

    static uint32_t even_ceil(double x) {
        uint32_t n = static_cast<uint32_t>(std::ceil(x - 1e-9));
        return (n % 2u) ? n + 1u : n;
    }

    // Barrel layer of flat rectangular modules tangent to radius R; staves
    // alternate radius by -1/+1 mm (phi overlaps), modules along z by 0/+0.5 mm.
    void barrel(double R, double half_z, double hu, double hv, double x0, double phi0) {
        const uint32_t li = static_cast<uint32_t>(layers.size());
        const uint32_t n_phi = even_ceil(kTwoPi * R / (2.0 * hu * 0.92));
        const uint32_t n_z = static_cast<uint32_t>(std::ceil(2.0 * half_z / (2.0 * hv * 0.96)));
        const double pitch = n_z > 1 ? (2.0 * half_z - 2.0 * hv) / (n_z - 1) : 0.0;
        const double z0 = -half_z + hv;
        ckf_layer L{};
        L.kind = CKF_LAYER_BARREL;
        L.first_surface = static_cast<uint32_t>(sfs.size());
        L.n_phi = n_phi;
        L.n_long = n_z;
        L.phi0 = static_cast<float>(phi0);
        L.long0 = static_cast<float>(z0);
        L.long_pitch = static_cast<float>(pitch > 0 ? pitch : 1.0);
        L.env_r[0] = static_cast<float>(R - 1.0 - 0.01);
        L.env_r[1] = static_cast<float>(std::sqrt((R + 1.5) * (R + 1.5) + hu * hu) + 0.01);
        L.env_z[0] = static_cast<float>(-half_z - 0.01);
        L.env_z[1] = static_cast<float>(half_z + 0.01);
        for (uint32_t ip = 0; ip < n_phi; ++ip) {
            const double phi = phi0 + kTwoPi * ip / n_phi;
            const double c = std::cos(phi), s = std::sin(phi);
            for (uint32_t iz = 0; iz < n_z; ++iz) {
                const double rm = R + ((ip % 2u) ? 1.0 : -1.0) + ((iz % 2u) ? 0.5 : 0.0);
                ckf_surface& S = new_surface(li, ip * n_z + iz);
                const double u[3] = {-s, c, 0.0};
                const double v[3] = {0.0, 0.0, 1.0};
                set_frame(S, rm * c, rm * s, z0 + iz * pitch, u, v);
                S.bound_kind = CKF_BOUND_RECTANGLE;
                S.bound[0] = static_cast<float>(hu);
                S.bound[1] = static_cast<float>(hv);
                S.thickness_x0 = static_cast<float>(x0);
            }
        }
        L.n_phi = n_phi;
        layers.push_back(L);
    }

    // One ring of trapezoid modules on a disk plane z_ring; modules alternate
    // z by -1/+1 mm; loc1 is radial, loc0 along phi.
    void ring(double z_ring, double r_in, double r_out, double x0, double phi0) {
        const uint32_t li = static_cast<uint32_t>(layers.size());
        const double rc = 0.5 * (r_in + r_out);
        const double h = 0.5 * (r_out - r_in);
        const uint32_t n_phi = even_ceil(kTwoPi * rc / 30.0);
        const double a = kPi * r_in / n_phi * 1.1;
        const double b = kPi * r_out / n_phi * 1.1;
        ckf_layer L{};
        L.kind = CKF_LAYER_RING;
        L.first_surface = static_cast<uint32_t>(sfs.size());
        L.n_phi = n_phi;
        L.n_long = 1u;
        L.phi0 = static_cast<float>(phi0);
        L.long0 = 0.f;
        L.long_pitch = 1.f;
        L.env_r[0] = static_cast<float>(r_in - 0.01);
        L.env_r[1] = static_cast<float>(std::sqrt(r_out * r_out + b * b) + 0.01);
        L.env_z[0] = static_cast<float>(z_ring - 1.0 - 0.01);
        L.env_z[1] = static_cast<float>(z_ring + 1.0 + 0.01);
        for (uint32_t ip = 0; ip < n_phi; ++ip) {
            const double phi = phi0 + kTwoPi * ip / n_phi;
            const double c = std::cos(phi), s = std::sin(phi);
            const double zm = z_ring + ((ip % 2u) ? 1.0 : -1.0);
            ckf_surface& S = new_surface(li, ip);
            const double u[3] = {s, -c, 0.0};
            const double v[3] = {c, s, 0.0};
            set_frame(S, rc * c, rc * s, zm, u, v);
            S.bound_kind = CKF_BOUND_TRAPEZOID;
            S.bound[0] = static_cast<float>(a);
            S.bound[1] = static_cast<float>(b);
            S.bound[2] = static_cast<float>(h);
            S.thickness_x0 = static_cast<float>(x0);
        }
        layers.push_back(L);
    }

    // Disk at |z| = z (sign included) with 4 rings covering r in [40, 316] mm.
    void disk(double z, double x0, double phi0) {
        for (int k = 0; k < 4; ++k) {
            const double r_in = 40.0 + 68.0 * k;
            const double dz = (k % 2) ? 2.0 : -2.0;
            ring(z + dz, r_in, r_in + 72.0, x0, phi0 + 0.05 * k);
        }
    }

    void telescope_plane(double x, double half, double x0) {
        const uint32_t li = static_cast<uint32_t>(layers.size());
        ckf_layer L{};
        L.kind = CKF_LAYER_PLANE;
        L.first_surface = static_cast<uint32_t>(sfs.size());
        L.n_phi = 1u;
        L.n_long = 1u;
        L.long_pitch = 1.f;
        L.env_r[0] = static_cast<float>(x - 0.01);
        L.env_r[1] = static_cast<float>(x + 0.01);
        L.env_z[0] = static_cast<float>(-half);
        L.env_z[1] = static_cast<float>(half);
        L.normal[0] = 1.f;
        ckf_surface& S = new_surface(li, 0u);
        const double u[3] = {0.0, 1.0, 0.0};
        const double v[3] = {0.0, 0.0, 1.0};
        set_frame(S, x, 0.0, 0.0, u, v);
        S.bound_kind = CKF_BOUND_RECTANGLE;
        S.bound[0] = static_cast<float>(half);
        S.bound[1] = static_cast<float>(half);
        S.thickness_x0 = static_cast<float>(x0);
        layers.push_back(L);
    }
};

template <typename V>
V* copy_out(const std::vector<V>& v) {
    V* p = static_cast<V*>(std::malloc(sizeof(V) * (v.empty() ? 1 : v.size())));
    if (!v.empty()) {
        std::memcpy(p, v.data(), sizeof(V) * v.size());
    }
    return p;
}

// ---------------------------------------------------------------------------
// Finite-solenoid field (Derby & Olbert, Am. J. Phys. 78 (2010) 229), with
// Bulirsch's generalised complete elliptic integral.
// ---------------------------------------------------------------------------
double cel(double kc, double p, double a, double b) {
    if (kc == 0.0) {
        return std::nan("");
    }
    const double errtol = 1e-12;
    double k = std::fabs(kc);
    double pp = p, aa = a, bb = b;
    double em = 1.0;
    if (p > 0.0) {
        pp = std::sqrt(p);
        bb = b / pp;
    } else {
        double f = kc * kc;
        double q = 1.0 - f;
        const double g = 1.0 - pp;
        f = f - pp;
        q = q * (bb - aa * pp);
        pp = std::sqrt(f / g);
        aa = (aa - bb) / g;
        bb = -q / (g * g * pp) + aa * pp;
    }
    double f = aa;
    aa = aa + bb / pp;
    double g = k / pp;
    bb = 2.0 * (bb + f * g);
    pp = g + pp;
    g = em;
    em = k + em;
    double kk = k;
    int guard = 0;
    while (std::fabs(g - k) > g * errtol && guard++ < 100) {
        k = 2.0 * std::sqrt(kk);
        kk = k * em;
        f = aa;
        aa = aa + bb / pp;
        g = kk / pp;
        bb = 2.0 * (bb + f * g);
        pp = g + pp;
        g = em;
        em = k + em;
    }
    return (kPi / 2.0) * (bb + aa * em) / (em * (em + pp));
}

// Field of an ideal solenoid with B0 = mu0 n I / pi = 1 (unnormalised).
void solenoid_unit(double radius, double half_len, double r, double z, double& br,
                   double& bz) {
    const double a = radius;
    const double zp = z + half_len;
    const double zm = z - half_len;
    const double ap = a / std::sqrt(zp * zp + (r + a) * (r + a));
    const double am = a / std::sqrt(zm * zm + (r + a) * (r + a));
    const double bp = zp / std::sqrt(zp * zp + (r + a) * (r + a));
    const double bm = zm / std::sqrt(zm * zm + (r + a) * (r + a));
    const double gamma = (a - r) / (a + r);
    const double kp = std::sqrt((zp * zp + (a - r) * (a - r)) / (zp * zp + (a + r) * (a + r)));
    const double km = std::sqrt((zm * zm + (a - r) * (a - r)) / (zm * zm + (a + r) * (a + r)));
    br = ap * cel(kp, 1.0, 1.0, -1.0) - am * cel(km, 1.0, 1.0, -1.0);
    bz = a / (a + r) *
         (bp * cel(kp, gamma * gamma, 1.0, gamma) - bm * cel(km, gamma * gamma, 1.0, gamma));
}

void solenoid_exact(double radius, double length, double b_centre, double r, double z,
                    double& br, double& bz) {
    double b0r = 0.0, b0z = 0.0;
    solenoid_unit(radius, 0.5 * length, 0.0, 0.0, b0r, b0z);
    double ur = 0.0, uz = 0.0;
    // the sheet itself (r == radius) is a discontinuity; the formula is
    // evaluated a hair inside
    const double rr = (std::fabs(r - radius) < 1e-9) ? r - 1e-6 : r;
    solenoid_unit(radius, 0.5 * length, rr, z, ur, uz);
    br = b_centre * ur / b0z;
    bz = b_centre * uz / b0z;
}

struct hit_record {
    uint32_t sf;
    double l0, l1;
    double phi, theta, qop, time;
};

}  // namespace

extern "C" {

int tamd_detector_build(uint32_t preset, float thickness_x0, uint32_t telescope_planes,
                        float telescope_spacing, tamd_detector* out) {
    if (out == nullptr) {
        return CKF_ERR_INVALID_ARGUMENT;
    }
    det_builder b;
    const double x0 = thickness_x0;
    if (preset == TAMD_DET_TELESCOPE) {
        if (telescope_planes == 0u || telescope_planes > 60u) {
            return CKF_ERR_INVALID_ARGUMENT;
        }
        for (uint32_t k = 0; k < telescope_planes; ++k) {
            b.telescope_plane(telescope_spacing * (k + 1u), 100000.0, x0);
        }
    } else if (preset <= TAMD_DET_BARREL5_DISK6) {
        std::vector<double> radii = {32.0, 72.0, 116.0, 172.0};
        std::vector<double> disks;
        if (preset == TAMD_DET_BARREL4_DISK4) {
            disks = {600.0, 900.0, 1200.0, 1500.0};
        } else if (preset == TAMD_DET_BARREL5_DISK6) {
            radii.push_back(260.0);
            disks = {600.0, 780.0, 960.0, 1140.0, 1320.0, 1500.0};
        }
        for (size_t i = 0; i < radii.size(); ++i) {
            b.barrel(radii[i], 500.0, 8.4, 36.0, x0, 0.1 * static_cast<double>(i));
        }
        for (auto it = disks.rbegin(); it != disks.rend(); ++it) {
            b.disk(-*it, x0, 0.0);
        }
        for (double z : disks) {
            b.disk(z, x0, 0.0);
        }
    } else {
        return CKF_ERR_INVALID_ARGUMENT;
    }
    if (b.layers.size() > 64u) {
        return CKF_ERR_UNSUPPORTED;
    }
    out->surfaces = copy_out(b.sfs);
    out->n_surfaces = static_cast<uint32_t>(b.sfs.size());
    out->layers = copy_out(b.layers);
    out->n_layers = static_cast<uint32_t>(b.layers.size());
    return CKF_OK;
}

void tamd_detector_free(tamd_detector* det) {
    if (det == nullptr) {
        return;
    }
    std::free(det->surfaces);
    std::free(det->layers);
    det->surfaces = nullptr;
    det->layers = nullptr;
    det->n_surfaces = det->n_layers = 0u;
}

double tamd_cel(double kc, double p, double a, double b) { return cel(kc, p, a, b); }

void tamd_solenoid_field_exact(float radius, float length, float b_centre, double r, double z,
                               double* br, double* bz) {
    solenoid_exact(radius, length, b_centre, r, z, *br, *bz);
}

int tamd_solenoid_map_build(float radius, float length, float b_centre, float r_max, float z_max,
                            float pitch, ckf_bfield* out, float** table) {
    if (out == nullptr || table == nullptr || !(pitch > 0.f) || !(r_max > 0.f) ||
        !(z_max > 0.f)) {
        return CKF_ERR_INVALID_ARGUMENT;
    }
    const uint32_t n_r = static_cast<uint32_t>(std::lround(r_max / pitch)) + 1u;
    const uint32_t n_z = static_cast<uint32_t>(std::lround(2.0 * z_max / pitch)) + 1u;
    float* t = static_cast<float*>(std::malloc(sizeof(float) * 2u * n_r * n_z));
    for (uint32_t iz = 0; iz < n_z; ++iz) {
        const double z = -static_cast<double>(z_max) + pitch * static_cast<double>(iz);
        for (uint32_t ir = 0; ir < n_r; ++ir) {
            const double r = pitch * static_cast<double>(ir);
            double br = 0.0, bz = 0.0;
            solenoid_exact(radius, length, b_centre, r, z, br, bz);
            t[2u * (iz * n_r + ir) + 0u] = static_cast<float>(br);
            t[2u * (iz * n_r + ir) + 1u] = static_cast<float>(bz);
        }
    }
    std::memset(out, 0, sizeof(*out));
    out->kind = CKF_BFIELD_RZ_MAP;
    out->map = t;
    out->n_r = n_r;
    out->n_z = n_z;
    out->r_min = 0.f;
    out->z_min = -z_max;
    out->dr = pitch;
    out->dz = pitch;
    *table = t;
    return CKF_OK;
}

void tamd_solenoid_map_free(float* table) { std::free(table); }

int tamd_event_generate(const ckf_detector* det, const ckf_bfield* field,
                        const tamd_event_spec* spec, tamd_event* out) {
    if (det == nullptr || field == nullptr || spec == nullptr || out == nullptr) {
        return CKF_ERR_INVALID_ARGUMENT;
    }
    std::memset(out, 0, sizeof(*out));
    rng R(spec->seed);

    ckf_finding_config fc;
    ckf_finding_config_default(&fc);
    fc.min_step_length_for_next_surface = 0.01f;
    fc.max_step_counts_for_next_surface = 2000u;
    fc.min_p = 0.f;
    fc.min_pT = 0.f;
    fc.ptc_mass = spec->mass;
    fc.ptc_charge = 1.f;
    fc.propagation.rk_error_tol = 1e-5f;
    fc.propagation.on_surface_tolerance = 1e-6f;
    fc.propagation.path_limit = 50000.f;

    tamd::navigator<double, false> nav{};
    nav.sfs = det->surfaces;
    nav.layers = det->layers;
    nav.n_layers = det->n_layers;
    nav.field = tamd::make_field_view<double>(*field, field->map);
    nav.cfg = tamd::make_nav_config<double>(fc);

    // dead modules
    std::vector<uint8_t> dead(det->n_surfaces, 0u);
    if (spec->dead_fraction > 0.f) {
        for (uint32_t i = 0; i < det->n_surfaces; ++i) {
            dead[i] = R.uniform() < spec->dead_fraction ? 1u : 0u;
        }
    }

    struct meas_tmp {
        uint32_t sf;
        float l0, l1, v0, v1, time;
        int32_t particle;
    };
    std::vector<meas_tmp> meas;
    std::vector<uint32_t> signal_per_sf(det->n_surfaces, 0u);
    struct seed_tmp {
        uint32_t particle;
        size_t meas_slot;
        hit_record hit;
    };
    std::vector<seed_tmp> seeds;

    const uint32_t np = spec->n_particles;
    std::vector<float> pv(3u * np), pm(3u * np), pq(np);
    std::vector<uint32_t> pn(np, 0u);
    const double mass = spec->mass;
    const uint32_t max_hits = spec->max_hits_per_particle ? spec->max_hits_per_particle : 64u;

    for (uint32_t ip = 0; ip < np; ++ip) {
        double vtx[3];
        for (int k = 0; k < 3; ++k) {
            vtx[k] = spec->vertex[k] + spec->vertex_sigma[k] * R.normal();
        }
        double pt = 0.0;
        if (spec->pt_distribution == TAMD_PT_UNIFORM) {
            pt = R.uniform(spec->pt_a, spec->pt_b);
        } else if (spec->pt_distribution == TAMD_PT_LOG_UNIFORM) {
            pt = std::exp(R.uniform(std::log(spec->pt_a), std::log(spec->pt_b)));
        } else {
            pt = spec->pt_a - spec->pt_b * std::log(R.uniform());
        }
        const double eta = R.uniform(spec->eta_min, spec->eta_max);
        const double phi = R.uniform(spec->phi_min, spec->phi_max);
        double charge = spec->pdg > 0 ? -1.0 : 1.0;
        if (std::abs(spec->pdg) == 211) {
            charge = spec->pdg > 0 ? 1.0 : -1.0;
        }
        if (spec->random_charge) {
            charge = R.uniform() < 0.5 ? -1.0 : 1.0;
        }
        double dir[3];
        double p = 0.0;
        if (spec->fixed_direction) {
            const double st = std::sin(spec->dir_theta), ct = std::cos(spec->dir_theta);
            dir[0] = std::cos(spec->dir_phi) * st;
            dir[1] = std::sin(spec->dir_phi) * st;
            dir[2] = ct;
            p = pt;  // the magnitude draw is |p| for a fixed-direction gun
        } else {
            const double ch = std::cosh(eta);
            dir[0] = std::cos(phi) / ch;
            dir[1] = std::sin(phi) / ch;
            dir[2] = std::tanh(eta);
            p = pt * ch;
        }
        for (int k = 0; k < 3; ++k) {
            pv[3u * ip + k] = static_cast<float>(vtx[k]);
            pm[3u * ip + k] = static_cast<float>(p * dir[k]);
        }
        pq[ip] = static_cast<float>(charge);

        nav.st.pos = {vtx[0], vtx[1], vtx[2]};
        nav.st.lo = {0.0, 0.0, 0.0};
        nav.st.dir = {dir[0], dir[1], dir[2]};
        nav.st.qop = charge / p;
        nav.st.time = 0.0;
        nav.cfg.abs_q = 1.0;
        int prev = -1;
        bool seeded = false;
        for (uint32_t nh = 0; nh < max_hits; ++nh) {
            uint32_t sf = 0u;
            const int status = nav.propagate(prev, sf);
            if (status != tamd::NAV_FOUND) {
                break;
            }
            const ckf_surface& S = det->surfaces[sf];
            const tamd::vec3<double> x = tamd::rel_to(nav.st, tamd::sf_center<double>(S));
            hit_record h;
            h.sf = sf;
            h.l0 = tamd::dot(x, tamd::sf_u<double>(S));
            h.l1 = tamd::dot(x, tamd::sf_v<double>(S));
            h.phi = std::atan2(nav.st.dir.y, nav.st.dir.x);
            h.theta = std::acos(std::max(-1.0, std::min(1.0, nav.st.dir.z)));
            h.qop = nav.st.qop;
            h.time = nav.st.time;
            if (!dead[sf]) {
                meas_tmp m;
                m.sf = sf;
                m.l0 = static_cast<float>(h.l0 + spec->smear[0] * R.normal());
                m.l1 = static_cast<float>(h.l1 + spec->smear[1] * R.normal());
                m.v0 = static_cast<float>(spec->smear[0] * spec->smear[0]);
                m.v1 = static_cast<float>(spec->smear[1] * spec->smear[1]);
                m.time = static_cast<float>(h.time);
                m.particle = static_cast<int32_t>(ip);
                meas.push_back(m);
                ++signal_per_sf[sf];
                ++pn[ip];
                if (!seeded) {
                    seeds.push_back({ip, meas.size() - 1u, h});
                    seeded = true;
                }
            }
            // multiple scattering in the module
            if (S.thickness_x0 > 0.f) {
                const tamd::vec3<double> w = tamd::sf_w<double>(S);
                const double cosi = std::fabs(tamd::dot(nav.st.dir, w));
                const double th0 = tamd::highland_theta0<double>(
                    S.thickness_x0 / std::max(cosi, 1e-3), nav.st.qop, mass, 1.0);
                const double a1 = th0 * R.normal();
                const double a2 = th0 * R.normal();
                const tamd::vec3<double> d = nav.st.dir;
                tamd::vec3<double> ref{0.0, 0.0, 1.0};
                if (std::fabs(d.z) > 0.9) {
                    ref = {1.0, 0.0, 0.0};
                }
                tamd::vec3<double> e1 = tamd::cross(d, ref);
                e1 = (1.0 / std::sqrt(tamd::dot(e1, e1))) * e1;
                const tamd::vec3<double> e2 = tamd::cross(d, e1);
                const double ang = std::sqrt(a1 * a1 + a2 * a2);
                if (ang > 0.0) {
                    const double cpsi = a1 / ang, spsi = a2 / ang;
                    tamd::vec3<double> nd = std::cos(ang) * d +
                                            std::sin(ang) * (cpsi * e1 + spsi * e2);
                    nav.st.dir = (1.0 / std::sqrt(tamd::dot(nd, nd))) * nd;
                }
            }
            prev = static_cast<int>(sf);
        }
    }

    // noise hits, uniform over the module area
    if (spec->noise_fraction > 0.f) {
        for (uint32_t sf = 0; sf < det->n_surfaces; ++sf) {
            if (signal_per_sf[sf] == 0u || dead[sf]) {
                continue;
            }
            const uint32_t nn = R.poisson(spec->noise_fraction * signal_per_sf[sf]);
            const ckf_surface& S = det->surfaces[sf];
            for (uint32_t k = 0; k < nn; ++k) {
                meas_tmp m;
                m.sf = sf;
                if (S.bound_kind == CKF_BOUND_RECTANGLE) {
                    m.l0 = static_cast<float>(R.uniform(-S.bound[0], S.bound[0]));
                    m.l1 = static_cast<float>(R.uniform(-S.bound[1], S.bound[1]));
                } else {
                    const double h = S.bound[2];
                    const double l1 = R.uniform(-h, h);
                    const double half = S.bound[0] + (S.bound[1] - S.bound[0]) * (l1 + h) / (2.0 * h);
                    m.l1 = static_cast<float>(l1);
                    m.l0 = static_cast<float>(R.uniform(-half, half));
                }
                m.v0 = static_cast<float>(spec->smear[0] * spec->smear[0]);
                m.v1 = static_cast<float>(spec->smear[1] * spec->smear[1]);
                m.time = 0.f;
                m.particle = -1;
                meas.push_back(m);
            }
        }
    }

    // sort by (surface, loc0, loc1, var0, var1)
    std::vector<uint32_t> order(meas.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const meas_tmp& A = meas[a];
        const meas_tmp& B = meas[b];
        const uint64_t ia = det->surfaces[A.sf].identifier;
        const uint64_t ib = det->surfaces[B.sf].identifier;
        if (ia != ib) return ia < ib;
        if (A.l0 != B.l0) return A.l0 < B.l0;
        if (A.l1 != B.l1) return A.l1 < B.l1;
        if (A.v0 != B.v0) return A.v0 < B.v0;
        return A.v1 < B.v1;
    });
    std::vector<uint32_t> slot_of(meas.size());
    for (uint32_t i = 0; i < order.size(); ++i) {
        slot_of[order[i]] = i;
    }

    const uint32_t nm = static_cast<uint32_t>(meas.size());
    out->n_meas = nm;
    out->local_position = static_cast<float*>(std::malloc(sizeof(float) * 2u * (nm + 1u)));
    out->local_variance = static_cast<float*>(std::malloc(sizeof(float) * 2u * (nm + 1u)));
    out->dimensions = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (nm + 1u)));
    out->time = static_cast<float*>(std::malloc(sizeof(float) * (nm + 1u)));
    out->diameter = static_cast<float*>(std::malloc(sizeof(float) * (nm + 1u)));
    out->identifier = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (nm + 1u)));
    out->surface_link = static_cast<uint64_t*>(std::malloc(sizeof(uint64_t) * (nm + 1u)));
    out->subspace = static_cast<uint8_t*>(std::malloc(2u * (nm + 1u)));
    out->cluster_index = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (nm + 1u)));
    out->meas_particle = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (nm + 1u)));
    for (uint32_t i = 0; i < nm; ++i) {
        const meas_tmp& m = meas[order[i]];
        out->local_position[2u * i] = m.l0;
        out->local_position[2u * i + 1u] = m.l1;
        out->local_variance[2u * i] = m.v0;
        out->local_variance[2u * i + 1u] = m.v1;
        out->dimensions[i] = 2u;
        out->time[i] = m.time;
        out->diameter[i] = 0.f;
        out->identifier[i] = i;
        out->surface_link[i] = det->surfaces[m.sf].identifier;
        out->subspace[2u * i] = 0u;
        out->subspace[2u * i + 1u] = 1u;
        out->cluster_index[i] = i;
        out->meas_particle[i] = m.particle;
    }

    const uint32_t ns = static_cast<uint32_t>(seeds.size());
    out->n_seeds = ns;
    out->seed_surface_link = static_cast<uint64_t*>(std::malloc(sizeof(uint64_t) * (ns + 1u)));
    out->seed_vector = static_cast<float*>(std::malloc(sizeof(float) * 6u * (ns + 1u)));
    out->seed_covariance = static_cast<float*>(std::malloc(sizeof(float) * 36u * (ns + 1u)));
    out->seed_particle = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (ns + 1u)));
    for (uint32_t i = 0; i < ns; ++i) {
        const seed_tmp& s = seeds[i];
        const meas_tmp& m = meas[s.meas_slot];
        out->seed_surface_link[i] = det->surfaces[m.sf].identifier;
        out->seed_particle[i] = s.particle;
        double phi = s.hit.phi + spec->seed_angle_sigma * R.normal();
        double theta = s.hit.theta + spec->seed_angle_sigma * R.normal();
        // wrap (traccc detail::wrap_phi_theta semantics)
        theta = std::fmod(theta, kTwoPi);
        if (theta < 0.0) theta += kTwoPi;
        if (theta > kPi) {
            phi += kPi;
            theta = kTwoPi - theta;
        }
        phi = phi - kTwoPi * std::floor((phi + kPi) / kTwoPi);
        const double qop = s.hit.qop * (1.0 + spec->seed_qop_rel_sigma * R.normal());
        float* v = out->seed_vector + 6u * i;
        v[0] = m.l0;
        v[1] = m.l1;
        v[2] = static_cast<float>(phi);
        v[3] = static_cast<float>(theta);
        v[4] = static_cast<float>(qop);
        v[5] = static_cast<float>(s.hit.time);
        float* c = out->seed_covariance + 36u * i;
        std::memset(c, 0, sizeof(float) * 36u);
        const double sa = spec->seed_angle_sigma;
        const double sq = spec->seed_qop_rel_sigma * std::fabs(qop);
        c[0 * 6 + 0] = m.v0;
        c[1 * 6 + 1] = m.v1;
        c[2 * 6 + 2] = static_cast<float>(sa * sa);
        c[3 * 6 + 3] = static_cast<float>(sa * sa);
        c[4 * 6 + 4] = static_cast<float>(sq * sq);
        c[5 * 6 + 5] = spec->seed_time_sigma * spec->seed_time_sigma;
    }
    (void)slot_of;

    out->n_particles = np;
    out->particle_vertex = copy_out(pv);
    out->particle_momentum = copy_out(pm);
    out->particle_charge = copy_out(pq);
    out->particle_n_meas = copy_out(pn);
    return CKF_OK;
}

void tamd_event_free(tamd_event* ev) {
    if (ev == nullptr) {
        return;
    }
    std::free(ev->local_position);
    std::free(ev->local_variance);
    std::free(ev->dimensions);
    std::free(ev->time);
    std::free(ev->diameter);
    std::free(ev->identifier);
    std::free(ev->surface_link);
    std::free(ev->subspace);
    std::free(ev->cluster_index);
    std::free(ev->meas_particle);
    std::free(ev->seed_surface_link);
    std::free(ev->seed_vector);
    std::free(ev->seed_covariance);
    std::free(ev->seed_particle);
    std::free(ev->particle_vertex);
    std::free(ev->particle_momentum);
    std::free(ev->particle_charge);
    std::free(ev->particle_n_meas);
    std::memset(ev, 0, sizeof(*ev));
}

}  // extern "C"
