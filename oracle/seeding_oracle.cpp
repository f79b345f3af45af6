// CPU oracle of the seeding chain -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/ and bench.py's cpu_baseline leg load it (liboracle_ckf.so), as
// the checker and the timed CPU baseline; the product never links it.  Built
// with -ffp-contract=off: every single-precision expression rounds as
// written, as in the device code, so both take the same cut decisions.
//
// This project's restatement of traccc's host seeding, in single precision
// like the reference (scalar = float):
//   * seedfinder_config::setup (core/include/traccc/seeding/detail/seeding_config.hpp:119-134)
//   * the spacepoint grid: circular phi axis and closed z axis
//     (core/include/traccc/seeding/grids/axis.hpp:37-420), their bin counts
//     (core/include/traccc/seeding/spacepoint_binning_helper.hpp:22-108), the
//     spacepoint selection (:110-127) and the binning in spacepoint order
//     (core/src/seeding/spacepoint_binning.cpp:35-53)
//   * per middle spacepoint, grid bin by grid bin (seed_finding.cpp:62-118):
//     bottom and top doublets over the phi x z neighbourhood in zone order
//     (doublet_finding.hpp:55-105, doublet_finding_helper.hpp:53-264),
//     triplets per bottom doublet with the compatible-seed weights
//     (triplet_finding.hpp:60-183, triplet_finding_helper.hpp:46-141),
//     then the filter: weight bonus, single-seed cut, sort, the per-middle
//     cut and at most maxSeedsPerSpM seeds (seed_filtering.cpp:28-122,
//     seed_selecting_helper.hpp:26-87, detail/triplet_sorter.hpp:35-70)
//   * track parameters from three spacepoints and the diagonal covariance
//     (track_params_estimation_helper.hpp:47-125, track_params_estimation.cpp:37-87)
// Deviation: triplets of one middle spacepoint that tie on both sort keys
// (weight, then the y^2 + z^2 sum of bottom and top) keep their finding
// order; the reference's std::sort leaves their order unspecified.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

constexpr float kPi = static_cast<float>(M_PI);

template <typename T>
T mn(T a, T b) {
    return (b < a) ? b : a;  // std::min
}

struct sp_t {
    float x, y, z, vr, vz;
    float radius() const { return std::sqrt(x * x + y * y); }
    float phi() const { return std::atan2(y, x); }
};

struct lin_circle {
    float Zo, cotTheta, iDeltaR, Er, U, V;
};

struct axes {
    uint32_t nphi, nz;
    float phi_min, phi_max, z_min, z_max;
    // circular bin
    uint32_t phi_bin(float v) const {
        const int ib = static_cast<int>((v - phi_min) / (phi_max - phi_min) * static_cast<float>(nphi));
        if (ib >= 0 && ib < static_cast<int>(nphi)) return static_cast<uint32_t>(ib);
        return ib < 0 ? nphi + static_cast<uint32_t>(ib) : static_cast<uint32_t>(ib) - nphi;
    }
    uint32_t phi_remap(uint32_t b, int s) const {
        const int o = static_cast<int>(b) + s;
        if (o >= 0 && o < static_cast<int>(nphi)) return static_cast<uint32_t>(o);
        if (o < 0) return static_cast<uint32_t>(static_cast<int>(nphi) + o);
        return static_cast<uint32_t>(o) - nphi;
    }
    std::vector<uint32_t> phi_zone(float v, const uint32_t nh[2]) const {
        const uint32_t g = phi_bin(v);
        const uint32_t lo = phi_remap(g, -static_cast<int>(nh[0]));
        const uint32_t hi = phi_remap(g, static_cast<int>(nh[1]));
        std::vector<uint32_t> s;
        if (lo < hi) {
            for (uint32_t b = lo; b <= hi; ++b) s.push_back(b);
            return s;
        }
        const uint32_t vl = nphi - lo + hi + 1u;
        uint32_t mi = 0, mo = 0;
        for (uint32_t k = 0; k < vl; ++k) {
            uint32_t n = lo + mi++;
            if (n > nphi - 1u) n = mo++;
            s.push_back(n);
        }
        return s;
    }
    // closed bin
    uint32_t z_bin(float v) const {
        const int ib = static_cast<int>((v - z_min) / (z_max - z_min) * static_cast<float>(nz));
        if (ib >= 0 && ib < static_cast<int>(nz)) return static_cast<uint32_t>(ib);
        return ib < 0 ? 0u : nz - 1u;
    }
    std::vector<uint32_t> z_zone(float v, const uint32_t nh[2]) const {
        const int ib = static_cast<int>((v - z_min) / (z_max - z_min) * static_cast<float>(nz));
        const int lo = ib - static_cast<int>(nh[0]);
        const int hi = ib + static_cast<int>(nh[1]);
        const uint32_t a = lo >= 0 ? static_cast<uint32_t>(lo) : 0u;
        const uint32_t b = hi < static_cast<int>(nz) ? static_cast<uint32_t>(hi) : nz - 1u;
        std::vector<uint32_t> s;
        for (uint32_t k = 0; k < b - a + 1u; ++k) s.push_back(a + k);
        return s;
    }
};

// spacepoint_binning_helper.hpp:22-108; false where the reference throws
bool make_axes(const ckf_spacepoint_grid_config& g, axes& A) {
    uint32_t phiBins;
    if (g.bFieldInZ == 0.f) {
        phiBins = 100u;
    } else {
        const float minHelixRadius = g.minPt / g.bFieldInZ;
        if (minHelixRadius < g.rMax / 2) return false;
        const float maxR2 = g.rMax * g.rMax;
        const float xOuter = maxR2 / (2 * minHelixRadius);
        const float yOuter = std::sqrt(maxR2 - xOuter * xOuter);
        const float outerAngle = std::atan(xOuter / yOuter);
        float innerAngle = 0;
        float rMin = g.rMax;
        if (g.rMax > g.deltaRMax) {
            rMin = g.rMax - g.deltaRMax;
            const float innerCircleR2 = (g.rMax - g.deltaRMax) * (g.rMax - g.deltaRMax);
            const float xInner = innerCircleR2 / (2 * minHelixRadius);
            const float yInner = std::sqrt(innerCircleR2 - xInner * xInner);
            innerAngle = std::atan(xInner / yInner);
        }
        const float deltaAngleWithMaxD0 =
            std::fabs(std::asin(g.impactMax / (rMin)) - std::asin(g.impactMax / g.rMax));
        const float deltaPhi = (outerAngle - innerAngle + deltaAngleWithMaxD0) /
                               static_cast<float>(g.phiBinDeflectionCoverage);
        if (deltaPhi <= 0.) return false;
        phiBins = static_cast<uint32_t>(std::llround(2 * M_PI / deltaPhi + 0.5));
    }
    const float zBinSize = g.cotThetaMax * g.deltaRMax;
    const uint32_t zBins =
        std::max(1u, static_cast<uint32_t>(std::floor((g.zMax - g.zMin) / zBinSize)));
    A = axes{phiBins, zBins, g.phiMin, g.phiMax, g.zMin, g.zMax};
    return true;
}

bool valid_sp(const ckf_seedfinder_config& c, const sp_t& s) {
    if (s.z > c.zMax || s.z < c.zMin) return false;
    const float phi = std::atan2(s.y, s.x);
    if (phi > c.phiMax || phi < c.phiMin) return false;
    const float bx = c.beamPos[0], by = c.beamPos[1];
    const float dx = s.x - bx, dy = s.y - by;
    const size_t nr = static_cast<size_t>(c.rMax + std::sqrt(bx * bx + by * by));
    return static_cast<size_t>(std::sqrt(dx * dx + dy * dy)) < nr;
}

// doublet_finding_helper.hpp:53-224 (bottom: other below the middle)
bool doublet_ok(bool bottom, const sp_t& m, const sp_t& o, const ckf_seedfinder_config& c) {
    float deltaR, cotTheta, zOrigin;
    if (bottom) {
        deltaR = m.radius() - o.radius();
        cotTheta = m.z - o.z;
        zOrigin = m.z * deltaR - m.radius() * cotTheta;
    } else {
        deltaR = o.radius() - m.radius();
        cotTheta = o.z - m.z;
        zOrigin = m.z * deltaR - m.radius() * cotTheta;
    }
    if ((deltaR >= c.deltaRMax) || (deltaR <= c.deltaRMin) ||
        (std::fabs(cotTheta) >= c.cotThetaMax * deltaR) || (zOrigin <= c.collisionRegionMin * deltaR) ||
        (zOrigin >= c.collisionRegionMax * deltaR) || std::fabs(cotTheta) >= c.deltaZMax) {
        return false;
    }
    const float midX = 0.5f * (m.x + o.x);
    const float midY = 0.5f * (m.y + o.y);
    const float slope = (o.y - m.y) / (o.x - m.x);
    const float deltaX = o.x - m.x;
    const float deltaY = o.y - m.y;
    const float deltaXY2 = deltaX * deltaX + deltaY * deltaY;
    const float sagitta = std::sqrt(c.minHelixRadius * c.minHelixRadius - deltaXY2 / 4.f);
    const float denom = std::sqrt((slope * slope + 1) / (slope * slope));
    const float cosA = 1.f / denom;
    const float sinA = -1.f / (slope * denom);
    const float mdx = sagitta * cosA;
    const float mdy = sagitta * sinA;
    const float mp1X = midX + mdx, mp2X = midX - mdx, mp1Y = midY + mdy, mp2Y = midY - mdy;
    const float mp1R2 = mp1X * mp1X + mp1Y * mp1Y;
    const float mp2R2 = mp2X * mp2X + mp2Y * mp2Y;
    if (mn(mp1R2, mp2R2) <= ((c.minHelixRadius - c.impactMax) * (c.minHelixRadius - c.impactMax))) {
        return false;
    }
    return true;
}

// doublet_finding_helper.hpp:226-264
lin_circle transform(bool bottom, const sp_t& m, const sp_t& o) {
    const float rM = m.radius();
    const float cosPhiM = m.x / rM;
    const float sinPhiM = m.y / rM;
    const float deltaX = o.x - m.x;
    const float deltaY = o.y - m.y;
    const float deltaZ = o.z - m.z;
    const float x = deltaX * cosPhiM + deltaY * sinPhiM;
    const float y = deltaY * cosPhiM - deltaX * sinPhiM;
    const float iDeltaR2 = 1.f / (deltaX * deltaX + deltaY * deltaY);
    const float iDeltaR = std::sqrt(iDeltaR2);
    float cot = deltaZ * iDeltaR;
    if (bottom) cot = -cot;
    lin_circle l;
    l.cotTheta = cot;
    l.Zo = m.z - rM * cot;
    l.iDeltaR = iDeltaR;
    l.U = x * iDeltaR2;
    l.V = y * iDeltaR2;
    l.Er = ((m.vz + o.vz) + (cot * cot) * (m.vr + o.vr)) * iDeltaR2;
    return l;
}

// triplet_finding_helper.hpp:46-141
bool triplet_ok(const sp_t& m, const lin_circle& lb, const lin_circle& lt, const ckf_seedfinder_config& c,
                float iSinTheta2, float scatteringInRegion2, float& curvature, float& impact) {
    const float error2 = lt.Er + lb.Er + 2.f * (lb.cotTheta * lt.cotTheta * m.vr + m.vz) * lb.iDeltaR * lt.iDeltaR;
    float deltaCotTheta = lb.cotTheta - lt.cotTheta;
    const float deltaCotTheta2 = deltaCotTheta * deltaCotTheta;
    float error = 0.f;
    float dCotThetaMinusError2 = 0.f;
    if (deltaCotTheta2 - error2 > 0) {
        deltaCotTheta = std::fabs(deltaCotTheta);
        error = std::sqrt(error2);
        dCotThetaMinusError2 = deltaCotTheta2 + error2 - 2.f * deltaCotTheta * error;
        if (dCotThetaMinusError2 > scatteringInRegion2) return false;
    }
    const float dU = lt.U - lb.U;
    if (dU == 0.f) return false;
    const float A = (lt.V - lb.V) / dU;
    const float S2 = 1.f + A * A;
    const float B = lb.V - A * lb.U;
    const float B2 = B * B;
    if (S2 < B2 * c.minHelixDiameter2) return false;
    const float iHelixDiameter2 = B2 / S2;
    float pT2scatter = 4.f * iHelixDiameter2 * c.pT2perRadius;
    const float pT = c.pTPerHelixRadius * std::sqrt(S2 / B2) / 2.f;
    if (pT > c.maxPtScattering) {
        const float pTscatter = c.highland / c.maxPtScattering;
        pT2scatter = pTscatter * pTscatter;
    }
    const float p2scatter = pT2scatter * iSinTheta2;
    if ((deltaCotTheta2 - error2 > 0.f) &&
        (dCotThetaMinusError2 > p2scatter * c.sigmaScattering * c.sigmaScattering)) {
        return false;
    }
    curvature = B / std::sqrt(S2);
    impact = std::fabs((A - B * m.radius()) * m.radius());
    return !(impact > c.impactMax);
}

struct triplet {
    uint32_t b, m, t;  // spacepoint indices
    float curvature, weight;
};

struct doublet {
    uint32_t sp;
    lin_circle lc;
};

}  // namespace

extern "C" void oracle_seedfinder_setup(ckf_seedfinder_config* c) {
    // seeding_config.hpp:119-134 (MeV = 0.001 GeV)
    c->highland = 13.6f * 0.001f * std::sqrt(c->radLengthPerSeed) *
                  (1.f + 0.038f * std::log(c->radLengthPerSeed));
    const float maxScatteringAngle = c->highland / c->minPt;
    c->maxScatteringAngle2 = maxScatteringAngle * maxScatteringAngle;
    c->pTPerHelixRadius = c->bFieldInZ;
    c->minHelixDiameter2 = std::pow(c->minPt * 2.f / c->pTPerHelixRadius, 2.f);
    c->minHelixRadius = std::sqrt(c->minHelixDiameter2) / 2.f;
    c->pT2perRadius = std::pow(c->highland / c->pTPerHelixRadius, 2.f);
}

extern "C" int oracle_seeding(const ckf_seedfinder_config* cfg, const ckf_spacepoint_grid_config* gcfg,
                              const ckf_seedfilter_config* fcfg, uint32_t n_sp, const float* global,
                              const float* var_r, const float* var_z, oracle_seeds* out) {
    if (!cfg || !gcfg || !fcfg || !out || (n_sp > 0 && !global)) return 1;
    std::memset(out, 0, sizeof(*out));
    const ckf_seedfinder_config& c = *cfg;
    const ckf_seedfilter_config& f = *fcfg;
    axes A{};
    if (!make_axes(*gcfg, A)) return 2;
    std::vector<sp_t> sps(n_sp);
    for (uint32_t i = 0; i < n_sp; ++i) {
        sps[i] = sp_t{global[3 * i], global[3 * i + 1], global[3 * i + 2], var_r ? var_r[i] : 0.f,
                      var_z ? var_z[i] : 0.f};
    }
    // the grid: bin = phi bin + n_phi * z bin, spacepoints in index order
    std::vector<std::vector<uint32_t>> bins(size_t(A.nphi) * A.nz);
    for (uint32_t i = 0; i < n_sp; ++i) {
        if (valid_sp(c, sps[i])) bins[A.phi_bin(sps[i].phi()) + A.nphi * A.z_bin(sps[i].z)].push_back(i);
    }
    std::vector<uint32_t> sb, sm, st;
    std::vector<float> sw;
    const float ss2 = c.sigmaScattering * c.sigmaScattering;
    for (size_t bi = 0; bi < bins.size(); ++bi) {
        for (const uint32_t mi : bins[bi]) {
            const sp_t& M = sps[mi];
            std::vector<doublet> bot, top;
            const std::vector<uint32_t> pz = A.phi_zone(M.phi(), c.neighbor_scope);
            const std::vector<uint32_t> zz = A.z_zone(M.z, c.neighbor_scope);
            for (int pass = 0; pass < 2; ++pass) {
                const bool bottom = pass == 0;
                for (const uint32_t pb : pz) {
                    for (const uint32_t zb : zz) {
                        for (const uint32_t oi : bins[pb + zb * A.nphi]) {
                            if (doublet_ok(bottom, M, sps[oi], c)) {
                                (bottom ? bot : top).push_back({oi, transform(bottom, M, sps[oi])});
                            }
                        }
                    }
                }
                if (bottom && bot.empty()) break;
            }
            if (bot.empty() || top.empty()) continue;
            std::vector<triplet> trips;
            for (const doublet& db : bot) {
                const float iSinTheta2 = 1.f + db.lc.cotTheta * db.lc.cotTheta;
                float scat2 = c.maxScatteringAngle2 * iSinTheta2;
                scat2 *= ss2;
                const size_t first = trips.size();
                for (const doublet& dt : top) {
                    float curv, imp;
                    if (!triplet_ok(M, db.lc, dt.lc, c, iSinTheta2, scat2, curv, imp)) continue;
                    trips.push_back({db.sp, mi, dt.sp, curv, -imp * f.impactWeightFactor});
                }
                // compatible-seed weights among this bottom's triplets
                for (size_t i = first; i < trips.size(); ++i) {
                    const float rt = sps[trips[i].t].radius();
                    std::vector<float> compat;
                    const float lo = trips[i].curvature - f.deltaInvHelixDiameter;
                    const float hi = trips[i].curvature + f.deltaInvHelixDiameter;
                    for (size_t j = first; j < trips.size(); ++j) {
                        if (i == j) continue;
                        const float ro = sps[trips[j].t].radius();
                        const float dr = rt - ro;
                        if (std::abs(dr) < f.deltaRMin) continue;
                        if (trips[j].curvature < lo) continue;
                        if (trips[j].curvature > hi) continue;
                        bool fresh = true;
                        for (const float prev : compat) {
                            if (std::abs(prev - ro) < f.deltaRMin) {
                                fresh = false;
                                break;
                            }
                        }
                        if (fresh) {
                            compat.push_back(ro);
                            trips[i].weight += f.compatSeedWeight;
                        }
                        if (compat.size() >= f.compatSeedLimit) break;
                    }
                }
            }
            // the filter
            std::vector<size_t> pass;
            for (size_t i = 0; i < trips.size(); ++i) {
                triplet& T = trips[i];
                const float rb = sps[T.b].radius(), rt = sps[T.t].radius();
                float w = 0;
                if (rb > f.good_spB_min_radius) w = f.good_spB_weight_increase;
                if (rt < f.good_spT_max_radius) w = f.good_spT_weight_increase;
                T.weight += w;
                if (rb > f.good_spB_min_radius && T.weight < f.good_spB_min_weight) continue;
                pass.push_back(i);
            }
            auto ysum = [&](const triplet& T) {
                const sp_t& B = sps[T.b];
                const sp_t& U = sps[T.t];
                return B.y * B.y + B.z * B.z + U.y * U.y + U.z * U.z;
            };
            std::stable_sort(pass.begin(), pass.end(), [&](size_t a, size_t b) {
                if (trips[a].weight != trips[b].weight) return trips[a].weight > trips[b].weight;
                return ysum(trips[a]) > ysum(trips[b]);
            });
            if (pass.empty()) continue;
            std::vector<size_t> fin{pass[0]};
            const size_t len = std::min(pass.size(), static_cast<size_t>(c.maxSeedsPerSpM));
            for (size_t i = 1; i < len; ++i) {
                const triplet& T = trips[pass[i]];
                if (T.weight > f.seed_min_weight || sps[T.b].radius() > f.spB_min_radius) fin.push_back(pass[i]);
            }
            for (size_t i = 0; i < fin.size() && i < c.maxSeedsPerSpM; ++i) {
                const triplet& T = trips[fin[i]];
                sb.push_back(T.b);
                sm.push_back(T.m);
                st.push_back(T.t);
                sw.push_back(T.weight);
            }
        }
    }
    const size_t n = sb.size();
    out->n_seeds = static_cast<uint32_t>(n);
    out->n_phi_bins = A.nphi;
    out->n_z_bins = A.nz;
    out->bottom = static_cast<uint32_t*>(std::malloc(4 * std::max<size_t>(n, 1)));
    out->middle = static_cast<uint32_t*>(std::malloc(4 * std::max<size_t>(n, 1)));
    out->top = static_cast<uint32_t*>(std::malloc(4 * std::max<size_t>(n, 1)));
    out->weight = static_cast<float*>(std::malloc(4 * std::max<size_t>(n, 1)));
    if (!out->bottom || !out->middle || !out->top || !out->weight) return 3;
    std::copy(sb.begin(), sb.end(), out->bottom);
    std::copy(sm.begin(), sm.end(), out->middle);
    std::copy(st.begin(), st.end(), out->top);
    std::copy(sw.begin(), sw.end(), out->weight);
    return 0;
}

extern "C" void oracle_seeds_free(oracle_seeds* s) {
    if (!s) return;
    std::free(s->bottom);
    std::free(s->middle);
    std::free(s->top);
    std::free(s->weight);
    std::memset(s, 0, sizeof(*s));
}

// track_params_estimation_helper.hpp:47-125 and track_params_estimation.cpp:60-87,
// in double precision (the checker of the device's single-precision result).
extern "C" int oracle_track_params(const ckf_track_params_estimation_config* cfg, const double b[3],
                                   const double p0[3], const double p1[3], const double p2[3],
                                   const double bottom_local[2], double vec[6], double cov_diag[6]) {
    if (!cfg) return 1;
    auto sub = [](const double* a, const double* c, double* o) {
        for (int k = 0; k < 3; ++k) o[k] = a[k] - c[k];
    };
    auto cross = [](const double* a, const double* c, double* o) {
        o[0] = a[1] * c[2] - a[2] * c[1];
        o[1] = a[2] * c[0] - a[0] * c[2];
        o[2] = a[0] * c[1] - a[1] * c[0];
    };
    auto norm = [](const double* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); };
    auto unit = [&](double* a) {
        const double n = norm(a);
        for (int k = 0; k < 3; ++k) a[k] /= n;
    };
    double rel[3], z[3] = {b[0], b[1], b[2]}, y[3], x[3];
    sub(p1, p0, rel);
    unit(z);
    cross(z, rel, y);
    unit(y);
    cross(y, z, x);
    auto to_local = [&](const double* p, double* l) {
        double d[3];
        sub(p, p0, d);
        l[0] = d[0] * x[0] + d[1] * x[1] + d[2] * x[2];
        l[1] = d[0] * y[0] + d[1] * y[1] + d[2] * y[2];
        l[2] = d[0] * z[0] + d[1] * z[1] + d[2] * z[2];
    };
    double l1[3], l2[3];
    to_local(p1, l1);
    to_local(p2, l2);
    const double d1 = l1[0] * l1[0] + l1[1] * l1[1], d2 = l2[0] * l2[0] + l2[1] * l2[1];
    const double u1 = l1[0] / d1, v1 = l1[1] / d1, u2 = l2[0] / d2, v2 = l2[1] / d2;
    const double A = (v2 - v1) / (u2 - u1);
    const double B = v2 - A * u2;
    const double R = -std::sqrt(1.0 + A * A) / (2.0 * B);
    const double perp2 = std::sqrt(l2[0] * l2[0] + l2[1] * l2[1]);
    const double invTanTheta = l2[2] / (2.0 * R * std::asin(perp2 / (2.0 * R)));
    double td[3] = {1.0, A, std::sqrt(1.0 + A * A) * invTanTheta};
    unit(td);
    const double dir[3] = {x[0] * td[0] + y[0] * td[1] + z[0] * td[2], x[1] * td[0] + y[1] * td[1] + z[1] * td[2],
                           x[2] * td[0] + y[2] * td[1] + z[2] * td[2]};
    vec[0] = bottom_local[0];
    vec[1] = bottom_local[1];
    vec[2] = std::atan2(dir[1], dir[0]);
    vec[3] = std::atan2(std::sqrt(dir[0] * dir[0] + dir[1] * dir[1]), dir[2]);
    const double qOverPt = 1.0 / (R * norm(b));
    vec[4] = qOverPt / std::sqrt(1.0 + invTanTheta * invTanTheta);
    vec[5] = 0.0;
    for (int j = 0; j < 6; ++j) {
        double var = double(cfg->initial_sigma[j]) * cfg->initial_sigma[j];
        if (j == 4) {
            const double vt = cov_diag[3];
            var += std::pow(cfg->initial_sigma_qopt * std::sin(vec[3]), 2.0);
            var += std::pow(cfg->initial_sigma_pt_rel * vec[4], 2.0);
            var += vt * std::pow(vec[4] / std::tan(vec[3]), 2.0);
        }
        var *= cfg->initial_inflation[j];
        cov_diag[j] = var;
    }
    return 0;
}
