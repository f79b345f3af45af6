// Triplet seeding, spacepoint formation and track-parameter estimation for
// MI355X (C ABI: include/traccc_amd/seeding.h).
//
// The reference seeds one middle spacepoint at a time, grid bin by grid bin
// (core/src/seeding/seed_finding.cpp:62-118): bottom and top doublets over
// the phi x z neighbourhood (doublet_finding.hpp:55-105), triplets per bottom
// doublet with compatible-seed weights (triplet_finding.hpp:60-183), and the
// per-middle filter (seed_filtering.cpp:28-122).  Here:
//   k_sd_prep       per spacepoint: radius, phi, the selection and the grid
//                   bin; key (bin << 32 | index)
//   radix sort      the keys: spacepoints in the reference's grid order
//                   (bin, then index), the middle order of the seeding
//   k_sd_doublets   one wave per middle spacepoint: the neighbourhood's
//                   candidates in zone order, 64 at a time; a ballot keeps
//                   the compatible ones in order (count pass, scan, fill pass)
//   k_sd_triplets   one wave per middle: per bottom doublet, the tops 64 at a
//                   time (count pass, scan, fill pass); every triplet records
//                   its bottom's segment
//   k_sd_weights    per triplet: the compatible-seed bonus over its
//                   segment, in the reference's order
//   k_sd_filter     one wave per middle: weight bonus, single-seed cut, the
//                   best maxSeedsPerSpM by (weight, y^2 + z^2 sum, finding
//                   order), the per-middle cut
//   k_sd_compact    the seeds in middle order
// Every single-precision expression rounds as written (no contraction), as
// in the oracle, so the cut decisions agree exactly.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../ckf/handles.hpp"
#include "../ckf/wave.hpp"
#include "../common/errors.hpp"
#include "../common/scan.hpp"
#include "../core/geometry.hpp"
#include "traccc_amd/seeding.h"

#pragma clang fp contract(off)

namespace tamd {
namespace {

constexpr uint32_t kSdBlock = 256;
constexpr uint32_t kSdWaves = kSdBlock / kWave;
constexpr uint32_t kNoSp = 0xffffffffu;

struct sd_axes {
    uint32_t nphi, nz;
    float phi_min, phi_max, z_min, z_max;
};

struct doublet_rec {
    uint32_t sp;
    float Zo, cot, iDR, Er, U, V;
};

struct triplet_rec {
    uint32_t b, t;
    float curv, w;
    uint32_t seg0, seg1;  // the triplets of the same bottom doublet
};

struct seed_rec {
    uint32_t b, m, t;
    float w;
};

struct sd_ctl {
    uint32_t n_valid;
    uint32_t n_bot, n_top, n_trip, n_seeds;
    uint32_t pad[3];
};

struct sd_params {
    ckf_seedfinder_config c;
    ckf_seedfilter_config f;
    sd_axes A;
    uint32_t n_sp;
    const float* xyz;
    const float* vr;  // may be null
    const float* vz;  // may be null
    float* rad;
    float* phi;
    uint64_t* keys;
    const uint64_t* sorted;
    uint32_t* bin_count;
    uint32_t* bin_off;
    uint32_t* cnt_b;  // per middle (grid order)
    uint32_t* off_b;
    uint32_t* cnt_t;
    uint32_t* off_t;
    doublet_rec* bot;
    doublet_rec* top;
    uint32_t* cnt_tr;
    uint32_t* off_tr;
    triplet_rec* trip;
    uint32_t* cnt_s;
    uint32_t* off_s;
    seed_rec* slots;
    seed_rec* seeds;
    sd_ctl* ctl;
};

__device__ inline float sp_vr(const sd_params& P, uint32_t i) { return P.vr ? P.vr[i] : 0.f; }
__device__ inline float sp_vz(const sd_params& P, uint32_t i) { return P.vz ? P.vz[i] : 0.f; }

// the reference's circular axis (core/include/traccc/seeding/grids/axis.hpp:227-420)
__device__ inline uint32_t phi_bin(const sd_axes& A, float v) {
    const int ib = static_cast<int>((v - A.phi_min) / (A.phi_max - A.phi_min) * static_cast<float>(A.nphi));
    if (ib >= 0 && ib < static_cast<int>(A.nphi)) return static_cast<uint32_t>(ib);
    return ib < 0 ? A.nphi + static_cast<uint32_t>(ib) : static_cast<uint32_t>(ib) - A.nphi;
}
__device__ inline uint32_t phi_remap(const sd_axes& A, uint32_t b, int s) {
    const int o = static_cast<int>(b) + s;
    if (o >= 0 && o < static_cast<int>(A.nphi)) return static_cast<uint32_t>(o);
    if (o < 0) return static_cast<uint32_t>(static_cast<int>(A.nphi) + o);
    return static_cast<uint32_t>(o) - A.nphi;
}
// the zone's bins: element q of count (a wrapping zone lists the bins past
// the last one from 0 again)
struct phi_zone {
    uint32_t lo, count, nphi;
    bool wraps;
    __device__ uint32_t at(uint32_t q) const {
        if (!wraps) return lo + q;
        const uint32_t n = lo + q;
        return n > nphi - 1u ? n - nphi : n;
    }
};
__device__ inline phi_zone make_phi_zone(const sd_axes& A, float v, const uint32_t* nh) {
    const uint32_t g = phi_bin(A, v);
    const uint32_t lo = phi_remap(A, g, -static_cast<int>(nh[0]));
    const uint32_t hi = phi_remap(A, g, static_cast<int>(nh[1]));
    phi_zone z;
    z.lo = lo;
    z.nphi = A.nphi;
    z.wraps = !(lo < hi);
    z.count = z.wraps ? A.nphi - lo + hi + 1u : hi - lo + 1u;
    return z;
}
// the reference's closed axis (axis.hpp:37-225)
__device__ inline uint32_t z_bin(const sd_axes& A, float v) {
    const int ib = static_cast<int>((v - A.z_min) / (A.z_max - A.z_min) * static_cast<float>(A.nz));
    if (ib >= 0 && ib < static_cast<int>(A.nz)) return static_cast<uint32_t>(ib);
    return ib < 0 ? 0u : A.nz - 1u;
}
__device__ inline void z_zone(const sd_axes& A, float v, const uint32_t* nh, uint32_t& a, uint32_t& count) {
    const int ib = static_cast<int>((v - A.z_min) / (A.z_max - A.z_min) * static_cast<float>(A.nz));
    const int lo = ib - static_cast<int>(nh[0]);
    const int hi = ib + static_cast<int>(nh[1]);
    a = lo >= 0 ? static_cast<uint32_t>(lo) : 0u;
    const uint32_t b = hi < static_cast<int>(A.nz) ? static_cast<uint32_t>(hi) : A.nz - 1u;
    count = b - a + 1u;
}

// doublet_finding_helper.hpp:53-224
__device__ inline bool doublet_ok(bool bottom, float mx, float my, float mz, float mr, float ox, float oy,
                                  float oz, float orr, const ckf_seedfinder_config& c) {
    float deltaR, cotTheta, zOrigin;
    if (bottom) {
        deltaR = mr - orr;
        cotTheta = mz - oz;
        zOrigin = mz * deltaR - mr * cotTheta;
    } else {
        deltaR = orr - mr;
        cotTheta = oz - mz;
        zOrigin = mz * deltaR - mr * cotTheta;
    }
    if ((deltaR >= c.deltaRMax) || (deltaR <= c.deltaRMin) || (fabsf(cotTheta) >= c.cotThetaMax * deltaR) ||
        (zOrigin <= c.collisionRegionMin * deltaR) || (zOrigin >= c.collisionRegionMax * deltaR) ||
        fabsf(cotTheta) >= c.deltaZMax) {
        return false;
    }
    const float midX = 0.5f * (mx + ox);
    const float midY = 0.5f * (my + oy);
    const float slope = (oy - my) / (ox - mx);
    const float deltaX = ox - mx;
    const float deltaY = oy - my;
    const float deltaXY2 = deltaX * deltaX + deltaY * deltaY;
    const float sagitta = sqrtf(c.minHelixRadius * c.minHelixRadius - deltaXY2 / 4.f);
    const float denom = sqrtf((slope * slope + 1) / (slope * slope));
    const float cosA = 1.f / denom;
    const float sinA = -1.f / (slope * denom);
    const float mdx = sagitta * cosA;
    const float mdy = sagitta * sinA;
    const float mp1X = midX + mdx, mp2X = midX - mdx, mp1Y = midY + mdy, mp2Y = midY - mdy;
    const float mp1R2 = mp1X * mp1X + mp1Y * mp1Y;
    const float mp2R2 = mp2X * mp2X + mp2Y * mp2Y;
    const float m = (mp2R2 < mp1R2) ? mp2R2 : mp1R2;  // std::min
    return !(m <= ((c.minHelixRadius - c.impactMax) * (c.minHelixRadius - c.impactMax)));
}

// doublet_finding_helper.hpp:226-264
__device__ inline doublet_rec transform(bool bottom, uint32_t osp, float mx, float my, float mz, float mr,
                                        float mvr, float mvz, float ox, float oy, float oz, float ovr,
                                        float ovz) {
    const float cosPhiM = mx / mr;
    const float sinPhiM = my / mr;
    const float deltaX = ox - mx;
    const float deltaY = oy - my;
    const float deltaZ = oz - mz;
    const float x = deltaX * cosPhiM + deltaY * sinPhiM;
    const float y = deltaY * cosPhiM - deltaX * sinPhiM;
    const float iDeltaR2 = 1.f / (deltaX * deltaX + deltaY * deltaY);
    const float iDeltaR = sqrtf(iDeltaR2);
    float cot = deltaZ * iDeltaR;
    if (bottom) cot = -cot;
    doublet_rec d;
    d.sp = osp;
    d.cot = cot;
    d.Zo = mz - mr * cot;
    d.iDR = iDeltaR;
    d.U = x * iDeltaR2;
    d.V = y * iDeltaR2;
    d.Er = ((mvz + ovz) + (cot * cot) * (mvr + ovr)) * iDeltaR2;
    return d;
}

// triplet_finding_helper.hpp:46-141
__device__ inline bool triplet_ok(float mr, float mvr, float mvz, const doublet_rec& lb, const doublet_rec& lt,
                                  const ckf_seedfinder_config& c, float iSinTheta2, float scat2, float& curvature,
                                  float& impact) {
    const float error2 = lt.Er + lb.Er + 2.f * (lb.cot * lt.cot * mvr + mvz) * lb.iDR * lt.iDR;
    float deltaCotTheta = lb.cot - lt.cot;
    const float deltaCotTheta2 = deltaCotTheta * deltaCotTheta;
    float dCotThetaMinusError2 = 0.f;
    if (deltaCotTheta2 - error2 > 0) {
        deltaCotTheta = fabsf(deltaCotTheta);
        const float error = sqrtf(error2);
        dCotThetaMinusError2 = deltaCotTheta2 + error2 - 2.f * deltaCotTheta * error;
        if (dCotThetaMinusError2 > scat2) return false;
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
    const float pT = c.pTPerHelixRadius * sqrtf(S2 / B2) / 2.f;
    if (pT > c.maxPtScattering) {
        const float pTscatter = c.highland / c.maxPtScattering;
        pT2scatter = pTscatter * pTscatter;
    }
    const float p2scatter = pT2scatter * iSinTheta2;
    if ((deltaCotTheta2 - error2 > 0.f) &&
        (dCotThetaMinusError2 > p2scatter * c.sigmaScattering * c.sigmaScattering)) {
        return false;
    }
    curvature = B / sqrtf(S2);
    impact = fabsf((A - B * mr) * mr);
    return !(impact > c.impactMax);
}

// spacepoint_binning_helper.hpp:110-127, spacepoint_binning.cpp:40-50
__global__ __launch_bounds__(kSdBlock) void k_sd_prep(sd_params P) {
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < P.n_sp; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        bool valid = false;
        if (i < P.n_sp) {
            const float x = P.xyz[3 * i], y = P.xyz[3 * i + 1], z = P.xyz[3 * i + 2];
            const float r = sqrtf(x * x + y * y);
            const float phi = atan2f(y, x);
            P.rad[i] = r;
            P.phi[i] = phi;
            const ckf_seedfinder_config& c = P.c;
            valid = !(z > c.zMax || z < c.zMin) && !(phi > c.phiMax || phi < c.phiMin);
            if (valid) {
                const float bx = c.beamPos[0], by = c.beamPos[1];
                const float dx = x - bx, dy = y - by;
                const size_t nr = static_cast<size_t>(c.rMax + sqrtf(bx * bx + by * by));
                valid = static_cast<size_t>(sqrtf(dx * dx + dy * dy)) < nr;
            }
            uint64_t key = ~0ull;
            if (valid) {
                const uint32_t bin = phi_bin(P.A, phi) + P.A.nphi * z_bin(P.A, z);
                key = (static_cast<uint64_t>(bin) << 32) | i;
                atomicAdd(&P.bin_count[bin], 1u);
            }
            P.keys[i] = key;
        }
        const uint64_t m = wave_ballot(valid);
        if (m != 0ull && lane() == static_cast<uint32_t>(__builtin_ctzll(m))) {
            atomicAdd(&P.ctl->n_valid, __popcll(m));
        }
    }
}

// One wave per middle spacepoint (grid order p): its bottom and top doublets.
template <bool kFill>
__global__ __launch_bounds__(kSdBlock) void k_sd_doublets(sd_params P) {
    const uint32_t n_mid = P.ctl->n_valid;
    const uint32_t l = lane();
    for (uint32_t p = blockIdx.x * kSdWaves + threadIdx.x / kWave; p < n_mid; p += gridDim.x * kSdWaves) {
        if (kFill && P.cnt_b[p] == 0u) {
            continue;  // no bottoms: nothing of this middle is kept
        }
        const uint32_t mi = static_cast<uint32_t>(P.sorted[p]);
        const float mx = P.xyz[3 * mi], my = P.xyz[3 * mi + 1], mz = P.xyz[3 * mi + 2];
        const float mr = P.rad[mi];
        const float mvr = sp_vr(P, mi), mvz = sp_vz(P, mi);
        const phi_zone pz = make_phi_zone(P.A, P.phi[mi], P.c.neighbor_scope);
        uint32_t za, zc;
        z_zone(P.A, mz, P.c.neighbor_scope, za, zc);
        uint32_t nb = 0u, nt = 0u;
        const uint32_t ob = kFill ? P.off_b[p] : 0u, ot = kFill ? P.off_t[p] : 0u;
        for (uint32_t qp = 0; qp < pz.count; ++qp) {
            const uint32_t pb = pz.at(qp);
            for (uint32_t qz = 0; qz < zc; ++qz) {
                const uint32_t bin = pb + (za + qz) * P.A.nphi;
                const uint32_t s0 = P.bin_off[bin], s1 = P.bin_off[bin + 1u];
                for (uint32_t c0 = s0; c0 < s1; c0 += kWave) {
                    const uint32_t k = c0 + l;
                    bool isb = false, ist = false;
                    uint32_t oi = 0u;
                    float ox = 0.f, oy = 0.f, oz = 0.f, orr = 0.f;
                    if (k < s1) {
                        oi = static_cast<uint32_t>(P.sorted[k]);
                        ox = P.xyz[3 * oi];
                        oy = P.xyz[3 * oi + 1];
                        oz = P.xyz[3 * oi + 2];
                        orr = P.rad[oi];
                        isb = doublet_ok(true, mx, my, mz, mr, ox, oy, oz, orr, P.c);
                        ist = doublet_ok(false, mx, my, mz, mr, ox, oy, oz, orr, P.c);
                    }
                    const uint64_t mb = wave_ballot(isb), mt = wave_ballot(ist);
                    if constexpr (kFill) {
                        if (isb) {
                            P.bot[ob + nb + prefix_count(mb)] =
                                transform(true, oi, mx, my, mz, mr, mvr, mvz, ox, oy, oz, sp_vr(P, oi), sp_vz(P, oi));
                        }
                        if (ist) {
                            P.top[ot + nt + prefix_count(mt)] =
                                transform(false, oi, mx, my, mz, mr, mvr, mvz, ox, oy, oz, sp_vr(P, oi), sp_vz(P, oi));
                        }
                    }
                    nb += __popcll(mb);
                    nt += __popcll(mt);
                }
            }
        }
        if (!kFill && l == 0u) {
            // a middle without bottoms makes no seed: its tops are not kept
            P.cnt_b[p] = nb;
            P.cnt_t[p] = nb > 0u ? nt : 0u;
        }
    }
}

// One wave per middle: per bottom doublet, its triplets with the tops.
template <bool kFill>
__global__ __launch_bounds__(kSdBlock) void k_sd_triplets(sd_params P) {
    const uint32_t n_mid = P.ctl->n_valid;
    const uint32_t l = lane();
    const float ss2 = P.c.sigmaScattering * P.c.sigmaScattering;
    for (uint32_t p = blockIdx.x * kSdWaves + threadIdx.x / kWave; p < n_mid; p += gridDim.x * kSdWaves) {
        const uint32_t nb = P.cnt_b[p], nt = P.cnt_t[p];
        if (nb == 0u || nt == 0u) {
            if (!kFill && l == 0u) P.cnt_tr[p] = 0u;
            continue;
        }
        const uint32_t mi = static_cast<uint32_t>(P.sorted[p]);
        const float mr = P.rad[mi], mvr = sp_vr(P, mi), mvz = sp_vz(P, mi);
        const doublet_rec* B = P.bot + P.off_b[p];
        const doublet_rec* T = P.top + P.off_t[p];
        uint32_t n = kFill ? P.off_tr[p] : 0u;
        for (uint32_t jb = 0; jb < nb; ++jb) {
            const doublet_rec lb = B[jb];
            const float iSinTheta2 = 1.f + lb.cot * lb.cot;
            float scat2 = P.c.maxScatteringAngle2 * iSinTheta2;
            scat2 *= ss2;
            // the segment of this bottom: counted first when filling
            uint32_t seg = 0u;
            if constexpr (kFill) {
                for (uint32_t c0 = 0; c0 < nt; c0 += kWave) {
                    bool ok = false;
                    if (c0 + l < nt) {
                        float cv, ip;
                        ok = triplet_ok(mr, mvr, mvz, lb, T[c0 + l], P.c, iSinTheta2, scat2, cv, ip);
                    }
                    seg += __popcll(wave_ballot(ok));
                }
            }
            uint32_t w = 0u;
            for (uint32_t c0 = 0; c0 < nt; c0 += kWave) {
                bool ok = false;
                float cv = 0.f, ip = 0.f;
                if (c0 + l < nt) ok = triplet_ok(mr, mvr, mvz, lb, T[c0 + l], P.c, iSinTheta2, scat2, cv, ip);
                const uint64_t m = wave_ballot(ok);
                if constexpr (kFill) {
                    if (ok) {
                        triplet_rec r;
                        r.b = lb.sp;
                        r.t = T[c0 + l].sp;
                        r.curv = cv;
                        r.w = -ip * P.f.impactWeightFactor;
                        r.seg0 = n;
                        r.seg1 = n + seg;
                        P.trip[n + w + prefix_count(m)] = r;
                    }
                }
                w += __popcll(m);
            }
            n += w;
        }
        if (!kFill && l == 0u) P.cnt_tr[p] = n;
    }
}

// triplet_finding.hpp:131-178: the compatible-seed bonus of each triplet
// over the triplets of its bottom doublet, in their order.
__global__ __launch_bounds__(kSdBlock) void k_sd_weights(sd_params P) {
    const uint32_t n = P.ctl->n_trip;
    const ckf_seedfilter_config& f = P.f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const triplet_rec ti = P.trip[i];
        const float rt = P.rad[ti.t];
        const float lo = ti.curv - f.deltaInvHelixDiameter;
        const float hi = ti.curv + f.deltaInvHelixDiameter;
        float compat[8];
        uint32_t nc = 0u;
        float w = ti.w;
        const uint32_t limit = f.compatSeedLimit < 8u ? f.compatSeedLimit : 8u;
        for (uint32_t j = ti.seg0; j < ti.seg1; ++j) {
            if (j == i) continue;
            const triplet_rec tj = P.trip[j];
            const float ro = P.rad[tj.t];
            const float dr = rt - ro;
            if (fabsf(dr) < f.deltaRMin) continue;
            if (tj.curv < lo) continue;
            if (tj.curv > hi) continue;
            bool fresh = true;
            for (uint32_t q = 0; q < nc; ++q) {
                if (fabsf(compat[q] - ro) < f.deltaRMin) {
                    fresh = false;
                    break;
                }
            }
            if (fresh) {
                compat[nc < 8u ? nc : 7u] = ro;
                ++nc;
                w += f.compatSeedWeight;
            }
            if (nc >= limit) break;
        }
        P.trip[i].w = w;
    }
}

struct seed_key {
    float w, ysum;
    uint32_t idx;
};
// a before b in the reference's sort (weight, then the y^2 + z^2 sum, both
// descending); the finding order on a full tie
__device__ inline bool before(const seed_key& a, const seed_key& b) {
    if (a.w != b.w) return a.w > b.w;
    if (a.ysum != b.ysum) return a.ysum > b.ysum;
    return a.idx < b.idx;
}

// seed_filtering.cpp:36-121 for one middle per wave.
__global__ __launch_bounds__(kSdBlock) void k_sd_filter(sd_params P) {
    const uint32_t n_mid = P.ctl->n_valid;
    const uint32_t l = lane();
    const ckf_seedfilter_config& f = P.f;
    const uint32_t max_s = P.c.maxSeedsPerSpM;
    for (uint32_t p = blockIdx.x * kSdWaves + threadIdx.x / kWave; p < n_mid; p += gridDim.x * kSdWaves) {
        const uint32_t t0 = P.off_tr[p], nt = P.cnt_tr[p];
        const uint32_t mi = static_cast<uint32_t>(P.sorted[p]);
        uint32_t out = 0u;
        seed_key last{0.f, 0.f, 0u};
        // the sorted passing triplets, one at a time: the best after `last`
        for (uint32_t s = 0; s < max_s && s < nt; ++s) {
            seed_key best{0.f, 0.f, kNoSp};
            for (uint32_t c0 = 0; c0 < nt; c0 += kWave) {
                const uint32_t k = c0 + l;
                if (k >= nt) continue;
                const triplet_rec r = P.trip[t0 + k];
                const float rb = P.rad[r.b], rt = P.rad[r.t];
                float bonus = 0;
                if (rb > f.good_spB_min_radius) bonus = f.good_spB_weight_increase;
                if (rt < f.good_spT_max_radius) bonus = f.good_spT_weight_increase;
                const float w = r.w + bonus;
                if (rb > f.good_spB_min_radius && w < f.good_spB_min_weight) continue;  // single-seed cut
                const float by = P.xyz[3 * r.b + 1], bz = P.xyz[3 * r.b + 2];
                const float ty = P.xyz[3 * r.t + 1], tz = P.xyz[3 * r.t + 2];
                const seed_key key{w, by * by + bz * bz + ty * ty + tz * tz, k};
                if (s > 0u && !before(last, key)) continue;  // already taken
                if (best.idx == kNoSp || before(key, best)) best = key;
            }
            // wave minimum under `before`
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                seed_key o;
                o.w = __shfl_xor(best.w, d, kWave);
                o.ysum = __shfl_xor(best.ysum, d, kWave);
                o.idx = __shfl_xor(best.idx, d, kWave);
                if (o.idx != kNoSp && (best.idx == kNoSp || before(o, best))) best = o;
            }
            if (best.idx == kNoSp) break;
            last = best;
            const triplet_rec r = P.trip[t0 + best.idx];
            // the first is always kept; the others need a heavy seed or a far bottom
            const bool keep = s == 0u || best.w > f.seed_min_weight || P.rad[r.b] > f.spB_min_radius;
            if (keep) {
                if (l == 0u) P.slots[static_cast<size_t>(p) * max_s + out] = seed_rec{r.b, mi, r.t, best.w};
                ++out;
            }
        }
        if (l == 0u) P.cnt_s[p] = out;
    }
}

__global__ __launch_bounds__(kSdBlock) void k_sd_compact(sd_params P) {
    const uint32_t n_mid = P.ctl->n_valid;
    const uint32_t max_s = P.c.maxSeedsPerSpM;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n_mid; p += gridDim.x * blockDim.x) {
        const uint32_t o = P.off_s[p];
        for (uint32_t s = 0; s < P.cnt_s[p]; ++s) P.seeds[o + s] = P.slots[static_cast<size_t>(p) * max_s + s];
    }
}

// core/include/traccc/seeding/impl/spacepoint_formation.ipp:20-48: a
// spacepoint per 2D measurement, global = centre + loc0 u + loc1 v.
__global__ __launch_bounds__(kSdBlock) void k_sd_sp_flag(const uint32_t* info, uint32_t m0, uint32_t n,
                                                         uint32_t* flag) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        flag[i] = (info[m0 + i] & 0xffu) == 2u ? 1u : 0u;
    }
}
__global__ __launch_bounds__(kSdBlock) void k_sd_sp_form(const ckf_surface* sfs, const float2* pos,
                                                         const uint32_t* meas_sf, uint32_t m0, uint32_t n,
                                                         const uint32_t* flag, const uint32_t* off,
                                                         float* xyz, uint32_t* midx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (flag[i] == 0u) continue;
        const ckf_surface& S = sfs[meas_sf[m0 + i]];
        const float2 lp = pos[m0 + i];
        const uint32_t o = off[i];
        for (int k = 0; k < 3; ++k) xyz[3 * o + k] = S.center[k] + lp.x * S.axis_u[k] + lp.y * S.axis_v[k];
        midx[o] = i;
    }
}

// track_params_estimation_helper.hpp:47-125 and track_params_estimation.cpp:60-87,
// one seed per thread, in single precision as the reference.
struct tp_params {
    ckf_track_params_estimation_config cfg;
    float b[3];
    uint32_t n;
    const float* xyz;
    const uint32_t* sb;
    const uint32_t* sm;
    const uint32_t* st;
    const float* bottom_local;  // [n][2]
    float* vec;                 // [n][6]
    float* cov;                 // [n][36]
};

__device__ inline void v_sub(const float* a, const float* c, float* o) {
    for (int k = 0; k < 3; ++k) o[k] = a[k] - c[k];
}
__device__ inline void v_cross(const float* a, const float* c, float* o) {
    o[0] = a[1] * c[2] - a[2] * c[1];
    o[1] = a[2] * c[0] - a[0] * c[2];
    o[2] = a[0] * c[1] - a[1] * c[0];
}
__device__ inline float v_norm(const float* a) { return sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
__device__ inline void v_unit(float* a) {
    const float n = v_norm(a);
    for (int k = 0; k < 3; ++k) a[k] = a[k] / n;
}

__global__ __launch_bounds__(kSdBlock) void k_track_params(tp_params P) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += gridDim.x * blockDim.x) {
        const float* p0 = P.xyz + 3u * P.sb[i];
        const float* p1 = P.xyz + 3u * P.sm[i];
        const float* p2 = P.xyz + 3u * P.st[i];
        // frame at the bottom spacepoint: z along the field, y normal to
        // the bottom -> middle chord
        float rel[3], z[3] = {P.b[0], P.b[1], P.b[2]}, y[3], x[3];
        v_sub(p1, p0, rel);
        v_unit(z);
        v_cross(z, rel, y);
        v_unit(y);
        v_cross(y, z, x);
        float l1[3], l2[3], d[3];
        v_sub(p1, p0, d);
        l1[0] = d[0] * x[0] + d[1] * x[1] + d[2] * x[2];
        l1[1] = d[0] * y[0] + d[1] * y[1] + d[2] * y[2];
        v_sub(p2, p0, d);
        l2[0] = d[0] * x[0] + d[1] * x[1] + d[2] * x[2];
        l2[1] = d[0] * y[0] + d[1] * y[1] + d[2] * y[2];
        l2[2] = d[0] * z[0] + d[1] * z[1] + d[2] * z[2];
        // conformal map: the circle through the three points is a line v = A u + B
        const float den1 = l1[0] * l1[0] + l1[1] * l1[1];
        const float den2 = l2[0] * l2[0] + l2[1] * l2[1];
        const float u1 = l1[0] / den1, v1 = l1[1] / den1, u2 = l2[0] / den2, v2 = l2[1] / den2;
        const float A = (v2 - v1) / (u2 - u1);
        const float B = v2 - A * u2;
        const float R = -sqrtf(1.f + A * A) / (2.f * B);
        const float perp2 = sqrtf(l2[0] * l2[0] + l2[1] * l2[1]);
        const float invTanTheta = l2[2] / (2.f * R * asinf(perp2 / (2.f * R)));
        float td[3] = {1.f, A, sqrtf(1.f + A * A) * invTanTheta};
        v_unit(td);
        const float dir[3] = {x[0] * td[0] + y[0] * td[1] + z[0] * td[2], x[1] * td[0] + y[1] * td[1] + z[1] * td[2],
                              x[2] * td[0] + y[2] * td[1] + z[2] * td[2]};
        float* v = P.vec + 6 * static_cast<size_t>(i);
        v[0] = P.bottom_local[2 * i];
        v[1] = P.bottom_local[2 * i + 1];
        v[2] = atan2f(dir[1], dir[0]);
        v[3] = atan2f(sqrtf(dir[0] * dir[0] + dir[1] * dir[1]), dir[2]);
        const float qOverPt = 1.f / (R * v_norm(P.b));
        v[4] = qOverPt / sqrtf(1.f + invTanTheta * invTanTheta);
        v[5] = 0.f;
        float* C = P.cov + 36 * static_cast<size_t>(i);
        for (int k = 0; k < 36; ++k) C[k] = 0.f;
        for (int j = 0; j < 6; ++j) {
            float var = P.cfg.initial_sigma[j] * P.cfg.initial_sigma[j];
            if (j == 4) {
                const float vt = C[3 * 6 + 3];
                const float a = P.cfg.initial_sigma_qopt * sinf(v[3]);
                const float b = P.cfg.initial_sigma_pt_rel * v[4];
                const float c = v[4] / tanf(v[3]);
                var += a * a;
                var += b * b;
                var += vt * (c * c);
            }
            var *= P.cfg.initial_inflation[j];
            C[j * 6 + j] = var;
        }
    }
}

uint32_t grid_for(uint64_t items, uint32_t per_block) {
    const uint64_t b = (items + per_block - 1u) / per_block;
    return static_cast<uint32_t>(std::max<uint64_t>(1u, std::min<uint64_t>(b, 65536u)));
}

// spacepoint_binning_helper.hpp:22-108; false where the reference throws
bool make_axes(const ckf_spacepoint_grid_config& g, sd_axes& A, std::string& why) {
    uint32_t phiBins;
    if (g.bFieldInZ == 0.f) {
        phiBins = 100u;
    } else {
        const float minHelixRadius = g.minPt / g.bFieldInZ;
        if (minHelixRadius < g.rMax / 2) {
            why = "the minimum helix radius (minPt / bFieldInZ) is smaller than rMax / 2";
            return false;
        }
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
        const float deltaAngleWithMaxD0 = std::fabs(std::asin(g.impactMax / (rMin)) - std::asin(g.impactMax / g.rMax));
        const float deltaPhi =
            (outerAngle - innerAngle + deltaAngleWithMaxD0) / static_cast<float>(g.phiBinDeflectionCoverage);
        if (deltaPhi <= 0.) {
            why = "the phi bin width is not positive";
            return false;
        }
        phiBins = static_cast<uint32_t>(std::llround(2 * M_PI / deltaPhi + 0.5));
    }
    const float zBinSize = g.cotThetaMax * g.deltaRMax;
    const uint32_t zBins = std::max(1u, static_cast<uint32_t>(std::floor((g.zMax - g.zMin) / zBinSize)));
    A = sd_axes{phiBins, zBins, g.phiMin, g.phiMax, g.zMin, g.zMax};
    return true;
}

}  // namespace
}  // namespace tamd

using tamd::ckf_error;
using tamd::guarded;
using tamd::hip_check;

struct ckf_seeding {
    ckf_seedfinder_config finder{};
    ckf_spacepoint_grid_config grid{};
    ckf_seedfilter_config filter{};
    tamd::sd_axes axes{};
    hipStream_t stream = nullptr;
    uint32_t n_seeds = 0;
    tamd::dbuf in_xyz, in_vr, in_vz, rad, phi, keys, sorted, sort_tmp, bin_count, bin_off, cnt_b, off_b,
        cnt_t, off_t, bot, top, cnt_tr, off_tr, trip, cnt_s, off_s, slots, seeds, sums, ctl;
    tamd::sd_ctl* h_ctl = nullptr;
    ~ckf_seeding() {
        tamd::dbuf* bs[] = {&in_xyz, &in_vr,  &in_vz,  &rad,   &phi,    &keys,  &sorted, &sort_tmp, &bin_count,
                            &bin_off, &cnt_b, &off_b,  &cnt_t, &off_t,  &bot,   &top,    &cnt_tr,   &off_tr,
                            &trip,   &cnt_s,  &off_s,  &slots, &seeds,  &sums,  &ctl};
        for (tamd::dbuf* b : bs) b->release();
        if (h_ctl != nullptr) (void)hipHostFree(h_ctl);
    }
};

namespace {

void read_ctl(ckf_seeding& S) {
    hip_check(hipMemcpyAsync(S.h_ctl, S.ctl.p, sizeof(tamd::sd_ctl), hipMemcpyDeviceToHost, S.stream), "readback");
    hip_check(hipStreamSynchronize(S.stream), "seeding");
}

void run_seeding(ckf_seeding& S, uint32_t n_sp, const float* d_xyz, const float* d_vr, const float* d_vz) {
    using namespace tamd;
    const hipStream_t s = S.stream;
    const uint32_t nbins = S.axes.nphi * S.axes.nz;
    const uint32_t N = std::max(n_sp, 1u);
    S.rad.ensure(4u * size_t(N));
    S.phi.ensure(4u * size_t(N));
    S.keys.ensure(8u * size_t(N));
    S.sorted.ensure(8u * size_t(N));
    S.bin_count.ensure(4u * size_t(nbins));
    S.bin_off.ensure(4u * (size_t(nbins) + 1u));
    S.sums.ensure(4u * size_t(scan_sums_size(std::max(N, nbins) + 1u)));
    S.ctl.ensure(sizeof(sd_ctl));
    if (S.h_ctl == nullptr) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&S.h_ctl), sizeof(sd_ctl)), "hipHostMalloc");
    }
    hip_check(hipMemsetAsync(S.ctl.p, 0, sizeof(sd_ctl), s), "memset");
    hip_check(hipMemsetAsync(S.bin_count.p, 0, 4u * size_t(nbins), s), "memset");
    sd_params P{};
    P.c = S.finder;
    P.f = S.filter;
    P.A = S.axes;
    P.n_sp = n_sp;
    P.xyz = d_xyz;
    P.vr = d_vr;
    P.vz = d_vz;
    P.rad = S.rad.as<float>();
    P.phi = S.phi.as<float>();
    P.keys = S.keys.as<uint64_t>();
    P.sorted = S.sorted.as<uint64_t>();
    P.bin_count = S.bin_count.as<uint32_t>();
    P.bin_off = S.bin_off.as<uint32_t>();
    P.ctl = S.ctl.as<sd_ctl>();
    if (n_sp > 0u) {
        hipLaunchKernelGGL(k_sd_prep, dim3(grid_for(n_sp, kSdBlock)), dim3(kSdBlock), 0, s, P);
        hip_check(hipGetLastError(), "prep");
        // (bin << 32 | index): the bits above the largest bin are zero
        int end_bit = 32;
        while (end_bit < 64 && (uint64_t(1) << (end_bit - 32)) <= nbins) ++end_bit;
        end_bit = 64;  // invalid spacepoints carry all-ones keys
        size_t tmp = 0;
        hip_check(rocprim::radix_sort_keys(nullptr, tmp, P.keys, const_cast<uint64_t*>(P.sorted), n_sp, 0,
                                           end_bit, s),
                  "sort size");
        S.sort_tmp.ensure(std::max<size_t>(tmp, 256));
        hip_check(rocprim::radix_sort_keys(S.sort_tmp.p, tmp, P.keys, const_cast<uint64_t*>(P.sorted), n_sp, 0,
                                           end_bit, s),
                  "sort");
    }
    exclusive_scan(P.bin_count, P.bin_off, S.sums.as<uint32_t>(), nbins, nullptr, s);
    read_ctl(S);
    const uint32_t nm = S.h_ctl->n_valid;
    const uint32_t M = std::max(nm, 1u);
    S.cnt_b.ensure(4u * size_t(M));
    S.off_b.ensure(4u * (size_t(M) + 1u));
    S.cnt_t.ensure(4u * size_t(M));
    S.off_t.ensure(4u * (size_t(M) + 1u));
    S.cnt_tr.ensure(4u * size_t(M));
    S.off_tr.ensure(4u * (size_t(M) + 1u));
    S.cnt_s.ensure(4u * size_t(M));
    S.off_s.ensure(4u * (size_t(M) + 1u));
    S.slots.ensure(sizeof(seed_rec) * size_t(M) * std::max(S.finder.maxSeedsPerSpM, 1u));
    S.sums.ensure(4u * size_t(scan_sums_size(M + 1u)));
    P.cnt_b = S.cnt_b.as<uint32_t>();
    P.off_b = S.off_b.as<uint32_t>();
    P.cnt_t = S.cnt_t.as<uint32_t>();
    P.off_t = S.off_t.as<uint32_t>();
    P.cnt_tr = S.cnt_tr.as<uint32_t>();
    P.off_tr = S.off_tr.as<uint32_t>();
    P.cnt_s = S.cnt_s.as<uint32_t>();
    P.off_s = S.off_s.as<uint32_t>();
    P.slots = S.slots.as<seed_rec>();
    const dim3 gw(grid_for(nm, kSdWaves));
    if (nm > 0u) {
        hipLaunchKernelGGL(k_sd_doublets<false>, gw, dim3(kSdBlock), 0, s, P);
        hip_check(hipGetLastError(), "doublet count");
    }
    exclusive_scan(P.cnt_b, P.off_b, S.sums.as<uint32_t>(), nm, &P.ctl->n_bot, s);
    exclusive_scan(P.cnt_t, P.off_t, S.sums.as<uint32_t>(), nm, &P.ctl->n_top, s);
    read_ctl(S);
    S.bot.ensure(sizeof(doublet_rec) * std::max<size_t>(S.h_ctl->n_bot, 1u));
    S.top.ensure(sizeof(doublet_rec) * std::max<size_t>(S.h_ctl->n_top, 1u));
    P.bot = S.bot.as<doublet_rec>();
    P.top = S.top.as<doublet_rec>();
    if (nm > 0u) {
        hipLaunchKernelGGL(k_sd_doublets<true>, gw, dim3(kSdBlock), 0, s, P);
        hipLaunchKernelGGL(k_sd_triplets<false>, gw, dim3(kSdBlock), 0, s, P);
        hip_check(hipGetLastError(), "triplet count");
    }
    exclusive_scan(P.cnt_tr, P.off_tr, S.sums.as<uint32_t>(), nm, &P.ctl->n_trip, s);
    read_ctl(S);
    S.trip.ensure(sizeof(triplet_rec) * std::max<size_t>(S.h_ctl->n_trip, 1u));
    P.trip = S.trip.as<triplet_rec>();
    if (nm > 0u) {
        hipLaunchKernelGGL(k_sd_triplets<true>, gw, dim3(kSdBlock), 0, s, P);
        hipLaunchKernelGGL(k_sd_weights, dim3(grid_for(std::max(S.h_ctl->n_trip, 1u), kSdBlock)), dim3(kSdBlock), 0,
                           s, P);
        hipLaunchKernelGGL(k_sd_filter, gw, dim3(kSdBlock), 0, s, P);
        hip_check(hipGetLastError(), "filter");
    }
    exclusive_scan(P.cnt_s, P.off_s, S.sums.as<uint32_t>(), nm, &P.ctl->n_seeds, s);
    read_ctl(S);
    S.n_seeds = S.h_ctl->n_seeds;
    S.seeds.ensure(sizeof(seed_rec) * std::max<size_t>(S.n_seeds, 1u));
    P.seeds = S.seeds.as<seed_rec>();
    if (nm > 0u) {
        hipLaunchKernelGGL(k_sd_compact, dim3(grid_for(nm, kSdBlock)), dim3(kSdBlock), 0, s, P);
        hip_check(hipGetLastError(), "compact");
    }
    hip_check(hipStreamSynchronize(s), "seeding");
}

}  // namespace

extern "C" {

ckf_status ckf_seeding_create(const ckf_seedfinder_config* finder, const ckf_spacepoint_grid_config* grid,
                              const ckf_seedfilter_config* filter, ckf_seeding** out) {
    return guarded([&] {
        if (!finder || !grid || !filter || !out) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        tamd::sd_axes A{};
        std::string why;
        if (!tamd::make_axes(*grid, A, why)) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, why);
        if (finder->maxSeedsPerSpM == 0u) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "maxSeedsPerSpM must be at least 1");
        }
        if (filter->compatSeedLimit > 8u) {
            throw ckf_error(CKF_ERR_UNSUPPORTED, "compatSeedLimit above 8 is not supported");
        }
        if (uint64_t(A.nphi) * A.nz > 0xfffffff0ull) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "too many grid bins");
        auto* s = new ckf_seeding();
        s->finder = *finder;
        s->grid = *grid;
        s->filter = *filter;
        s->axes = A;
        *out = s;
    });
}

ckf_status ckf_seeding_destroy(ckf_seeding* s) {
    return guarded([&] { delete s; });
}

ckf_status ckf_seeding_set_stream(ckf_seeding* s, void* hip_stream) {
    return guarded([&] {
        if (s == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null seeding");
        s->stream = static_cast<hipStream_t>(hip_stream);
    });
}

ckf_status ckf_seeding_run(ckf_seeding* s, const ckf_spacepoints* sp, uint32_t* n_seeds) {
    return guarded([&] {
        if (s == nullptr || sp == nullptr || (sp->size > 0u && sp->global == nullptr)) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        const uint32_t n = sp->size;
        const size_t N = std::max<size_t>(n, 1u);
        s->in_xyz.ensure(12u * N);
        if (n > 0u) {
            hip_check(hipMemcpyAsync(s->in_xyz.p, sp->global, 12u * size_t(n), hipMemcpyHostToDevice, s->stream),
                      "upload");
        }
        const float* vr = nullptr;
        const float* vz = nullptr;
        if (sp->radius_variance != nullptr && n > 0u) {
            s->in_vr.ensure(4u * N);
            hip_check(hipMemcpyAsync(s->in_vr.p, sp->radius_variance, 4u * size_t(n), hipMemcpyHostToDevice,
                                     s->stream), "upload");
            vr = s->in_vr.as<float>();
        }
        if (sp->z_variance != nullptr && n > 0u) {
            s->in_vz.ensure(4u * N);
            hip_check(hipMemcpyAsync(s->in_vz.p, sp->z_variance, 4u * size_t(n), hipMemcpyHostToDevice, s->stream),
                      "upload");
            vz = s->in_vz.as<float>();
        }
        run_seeding(*s, n, s->in_xyz.as<float>(), vr, vz);
        if (n_seeds) *n_seeds = s->n_seeds;
    });
}

ckf_status ckf_seeding_run_device(ckf_seeding* s, const ckf_spacepoints* sp, uint32_t* n_seeds) {
    return guarded([&] {
        if (s == nullptr || sp == nullptr || (sp->size > 0u && sp->global == nullptr)) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        run_seeding(*s, sp->size, sp->global, sp->radius_variance, sp->z_variance);
        if (n_seeds) *n_seeds = s->n_seeds;
    });
}

ckf_status ckf_seeding_result_size(const ckf_seeding* s, uint32_t* n_seeds) {
    return guarded([&] {
        if (s == nullptr || n_seeds == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        *n_seeds = s->n_seeds;
    });
}

ckf_status ckf_seeding_result_download(const ckf_seeding* s, ckf_seeds* out) {
    return guarded([&] {
        if (s == nullptr || out == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        if (out->size < s->n_seeds) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "output smaller than the seeds");
        std::vector<tamd::seed_rec> h(s->n_seeds);
        if (s->n_seeds > 0u) {
            hip_check(hipMemcpy(h.data(), s->seeds.p, sizeof(tamd::seed_rec) * h.size(), hipMemcpyDeviceToHost),
                      "download");
        }
        for (uint32_t i = 0; i < s->n_seeds; ++i) {
            if (out->bottom) out->bottom[i] = h[i].b;
            if (out->middle) out->middle[i] = h[i].m;
            if (out->top) out->top[i] = h[i].t;
            if (out->weight) out->weight[i] = h[i].w;
        }
    });
}

ckf_status ckf_spacepoint_formation(const ckf_device_detector* det, const ckf_device_event* ev, uint32_t event,
                                    float* global, uint32_t* measurement_index, uint32_t* n) {
    return guarded([&] {
        if (det == nullptr || ev == nullptr || n == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        if (event >= ev->n_events) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "event index out of range");
        if (ev->n_surfaces != det->h_sfs.size()) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "events and detector do not match");
        }
        const uint32_t m0 = ev->meas_off[event], nm = ev->meas_off[event + 1] - m0;
        const hipStream_t s = nullptr;
        tamd::dbuf flag, off, sums, xyz, midx;
        const uint32_t N = std::max(nm, 1u);
        flag.ensure(4u * size_t(N));
        off.ensure(4u * (size_t(N) + 1u));
        sums.ensure(4u * size_t(tamd::scan_sums_size(N)));
        uint32_t* d_total = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void**>(&d_total), 4), "hipMalloc");
        try {
            if (nm > 0u) {
                hipLaunchKernelGGL(tamd::k_sd_sp_flag, dim3(tamd::grid_for(nm, tamd::kSdBlock)), dim3(tamd::kSdBlock),
                                   0, s, ev->d_info, m0, nm, flag.as<uint32_t>());
                hip_check(hipGetLastError(), "spacepoint flags");
            }
            tamd::exclusive_scan(flag.as<uint32_t>(), off.as<uint32_t>(), sums.as<uint32_t>(), nm, d_total, s);
            uint32_t total = 0;
            hip_check(hipMemcpy(&total, d_total, 4, hipMemcpyDeviceToHost), "readback");
            *n = total;
            if ((global != nullptr || measurement_index != nullptr) && total > 0u) {
                xyz.ensure(12u * size_t(total));
                midx.ensure(4u * size_t(total));
                hipLaunchKernelGGL(tamd::k_sd_sp_form, dim3(tamd::grid_for(nm, tamd::kSdBlock)), dim3(tamd::kSdBlock), 0,
                                   s, det->d_sfs, ev->d_pos, ev->d_meas_sf, m0, nm, flag.as<uint32_t>(),
                                   off.as<uint32_t>(), xyz.as<float>(), midx.as<uint32_t>());
                hip_check(hipGetLastError(), "spacepoints");
                if (global) hip_check(hipMemcpy(global, xyz.p, 12u * size_t(total), hipMemcpyDeviceToHost), "download");
                if (measurement_index) {
                    hip_check(hipMemcpy(measurement_index, midx.p, 4u * size_t(total), hipMemcpyDeviceToHost),
                              "download");
                }
            }
        } catch (...) {
            (void)hipFree(d_total);
            for (tamd::dbuf* b : {&flag, &off, &sums, &xyz, &midx}) b->release();
            throw;
        }
        (void)hipFree(d_total);
        for (tamd::dbuf* b : {&flag, &off, &sums, &xyz, &midx}) b->release();
    });
}

ckf_status ckf_estimate_track_params(const ckf_track_params_estimation_config* cfg, const ckf_spacepoints* sp,
                                     const ckf_seeds* seeds, const ckf_measurements* meas, const float bfield[3],
                                     uint64_t* surface_link, float* vector, float* covariance) {
    return guarded([&] {
        if (!cfg || !sp || !seeds || !meas || !bfield) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        const uint32_t n = seeds->size;
        if (n == 0u) return;
        if (!seeds->bottom || !seeds->middle || !seeds->top || !sp->global || !sp->measurement_index_1 ||
            !meas->local_position || !meas->surface_link) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null seed, spacepoint or measurement arrays");
        }
        // the bottom spacepoint's measurement: local position and surface (host gather)
        std::vector<float> bl(2u * size_t(n));
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t b = seeds->bottom[i];
            if (seeds->bottom[i] >= sp->size || seeds->middle[i] >= sp->size || seeds->top[i] >= sp->size) {
                throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "seed spacepoint index out of range");
            }
            const uint32_t m = sp->measurement_index_1[b];
            if (m >= meas->size) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "spacepoint measurement index out of range");
            bl[2 * i] = meas->local_position[2 * m];
            bl[2 * i + 1] = meas->local_position[2 * m + 1];
            if (surface_link) surface_link[i] = meas->surface_link[m];
        }
        tamd::dbuf d_xyz, d_idx, d_bl, d_vec, d_cov;
        d_xyz.ensure(12u * size_t(sp->size));
        d_idx.ensure(12u * size_t(n));
        d_bl.ensure(8u * size_t(n));
        d_vec.ensure(24u * size_t(n));
        d_cov.ensure(144u * size_t(n));
        try {
            hip_check(hipMemcpy(d_xyz.p, sp->global, 12u * size_t(sp->size), hipMemcpyHostToDevice), "upload");
            uint32_t* di = d_idx.as<uint32_t>();
            hip_check(hipMemcpy(di, seeds->bottom, 4u * size_t(n), hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(di + n, seeds->middle, 4u * size_t(n), hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(di + 2u * n, seeds->top, 4u * size_t(n), hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(d_bl.p, bl.data(), 8u * size_t(n), hipMemcpyHostToDevice), "upload");
            tamd::tp_params P{};
            P.cfg = *cfg;
            for (int k = 0; k < 3; ++k) P.b[k] = bfield[k];
            P.n = n;
            P.xyz = d_xyz.as<float>();
            P.sb = di;
            P.sm = di + n;
            P.st = di + 2u * n;
            P.bottom_local = d_bl.as<float>();
            P.vec = d_vec.as<float>();
            P.cov = d_cov.as<float>();
            hipLaunchKernelGGL(tamd::k_track_params, dim3(tamd::grid_for(n, tamd::kSdBlock)), dim3(tamd::kSdBlock), 0,
                               nullptr, P);
            hip_check(hipGetLastError(), "track parameters");
            if (vector) hip_check(hipMemcpy(vector, d_vec.p, 24u * size_t(n), hipMemcpyDeviceToHost), "download");
            if (covariance) hip_check(hipMemcpy(covariance, d_cov.p, 144u * size_t(n), hipMemcpyDeviceToHost), "download");
        } catch (...) {
            for (tamd::dbuf* b : {&d_xyz, &d_idx, &d_bl, &d_vec, &d_cov}) b->release();
            throw;
        }
        for (tamd::dbuf* b : {&d_xyz, &d_idx, &d_bl, &d_vec, &d_cov}) b->release();
    });
}

}  // extern "C"
