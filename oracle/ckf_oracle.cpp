// CPU oracle for the CKF track-finding path -- TEST INFRASTRUCTURE ONLY.
//
// Restates traccc's host CKF (core/include/traccc/finding/details/
// combinatorial_kalman_filter.hpp:59-718) step for step, with dense-matrix
// Kalman algebra as in the reference (measurement_selector.hpp:152-212,
// gain_matrix_updater.hpp:69-181) and this project's propagation model
// (RK4-Nystroem + semi-analytic transport + layer-window navigation) written
// independently of the product's kernels, in dense 8x8 free-parameter form.
// Templated on the scalar: double = oracle, float = timed CPU baseline.
//
// Deliberate deviation from the reference, applied identically in the
// product: the per-surface branch sort breaks chi2 ties by the lower
// measurement index (traccc host uses std::sort, whose tie order is
// unspecified; the traccc device kernel uses the same lower-index rule,
// device/common/include/traccc/finding/device/impl/find_tracks.ipp:474-477).
#include "oracle.h"
#include "traccc_amd/fitting.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <thread>
#include <utility>
#include <vector>

namespace oracle {

constexpr double kPi = 3.14159265358979323846;
constexpr double kUnitT = 0.000299792458;
constexpr double kC = 299.792458;

// ---------------------------------------------------------------------------
// Dense matrices
// ---------------------------------------------------------------------------
template <typename S, int R, int C>
struct mat {
    S a[R][C];
    static mat zero() {
        mat m;
        for (int i = 0; i < R; ++i)
            for (int j = 0; j < C; ++j) m.a[i][j] = S(0);
        return m;
    }
    static mat identity() {
        mat m = zero();
        for (int i = 0; i < (R < C ? R : C); ++i) m.a[i][i] = S(1);
        return m;
    }
    S& operator()(int i, int j) { return a[i][j]; }
    const S& operator()(int i, int j) const { return a[i][j]; }
};

template <typename S, int R, int K, int C>
mat<S, R, C> operator*(const mat<S, R, K>& x, const mat<S, K, C>& y) {
    mat<S, R, C> m;
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < C; ++j) {
            S acc = S(0);
            for (int k = 0; k < K; ++k) acc += x.a[i][k] * y.a[k][j];
            m.a[i][j] = acc;
        }
    return m;
}
template <typename S, int R, int C>
mat<S, R, C> operator+(const mat<S, R, C>& x, const mat<S, R, C>& y) {
    mat<S, R, C> m;
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < C; ++j) m.a[i][j] = x.a[i][j] + y.a[i][j];
    return m;
}
template <typename S, int R, int C>
mat<S, R, C> operator-(const mat<S, R, C>& x, const mat<S, R, C>& y) {
    mat<S, R, C> m;
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < C; ++j) m.a[i][j] = x.a[i][j] - y.a[i][j];
    return m;
}
template <typename S, int R, int C>
mat<S, C, R> transpose(const mat<S, R, C>& x) {
    mat<S, C, R> m;
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < C; ++j) m.a[j][i] = x.a[i][j];
    return m;
}

template <typename S>
using v3 = mat<S, 3, 1>;

template <typename S>
v3<S> mk3(S x, S y, S z) {
    v3<S> v;
    v.a[0][0] = x;
    v.a[1][0] = y;
    v.a[2][0] = z;
    return v;
}
template <typename S>
S dot3(const v3<S>& a, const v3<S>& b) {
    return a.a[0][0] * b.a[0][0] + a.a[1][0] * b.a[1][0] + a.a[2][0] * b.a[2][0];
}
template <typename S>
v3<S> cross3(const v3<S>& a, const v3<S>& b) {
    return mk3<S>(a.a[1][0] * b.a[2][0] - a.a[2][0] * b.a[1][0],
                  a.a[2][0] * b.a[0][0] - a.a[0][0] * b.a[2][0],
                  a.a[0][0] * b.a[1][0] - a.a[1][0] * b.a[0][0]);
}
template <typename S>
v3<S> scale3(S s, const v3<S>& a) {
    return mk3<S>(s * a.a[0][0], s * a.a[1][0], s * a.a[2][0]);
}
template <typename S>
v3<S> fl3(const float* f) {
    return mk3<S>(static_cast<S>(f[0]), static_cast<S>(f[1]), static_cast<S>(f[2]));
}

// ---------------------------------------------------------------------------
// Bound / free parameter containers
// ---------------------------------------------------------------------------
template <typename S>
struct bound_state {
    uint32_t surface;
    mat<S, 6, 1> vec;
    mat<S, 6, 6> cov;
};

// ---------------------------------------------------------------------------
// Field: constant or bilinear (r, z) table, clamped at the grid edge.
// ---------------------------------------------------------------------------
template <typename S>
v3<S> bfield(const ckf_bfield& f, const v3<S>& p) {
    if (f.kind == CKF_BFIELD_CONSTANT) {
        return mk3<S>(static_cast<S>(f.b[0] * kUnitT), static_cast<S>(f.b[1] * kUnitT),
                      static_cast<S>(f.b[2] * kUnitT));
    }
    const S x = p.a[0][0], y = p.a[1][0], z = p.a[2][0];
    const S r = std::sqrt(x * x + y * y);
    S gr = (r - static_cast<S>(f.r_min)) * static_cast<S>(1.0 / f.dr);
    S gz = (z - static_cast<S>(f.z_min)) * static_cast<S>(1.0 / f.dz);
    gr = std::min(std::max(gr, S(0)), static_cast<S>(f.n_r - 1));
    gz = std::min(std::max(gz, S(0)), static_cast<S>(f.n_z - 1));
    const int ir = std::min(static_cast<int>(gr), static_cast<int>(f.n_r) - 2);
    const int iz = std::min(static_cast<int>(gz), static_cast<int>(f.n_z) - 2);
    const S fr = gr - static_cast<S>(ir);
    const S fz = gz - static_cast<S>(iz);
    auto at = [&](int jz, int jr, int c) {
        return static_cast<S>(f.map[2 * (jz * static_cast<int>(f.n_r) + jr) + c]);
    };
    S b[2];
    for (int c = 0; c < 2; ++c) {
        b[c] = (S(1) - fr) * (S(1) - fz) * at(iz, ir, c) + fr * (S(1) - fz) * at(iz, ir + 1, c) +
               (S(1) - fr) * fz * at(iz + 1, ir, c) + fr * fz * at(iz + 1, ir + 1, c);
        b[c] *= static_cast<S>(kUnitT);
    }
    if (r > S(0)) {
        return mk3<S>(b[0] * x / r, b[0] * y / r, b[1]);
    }
    return mk3<S>(S(0), S(0), b[1]);
}

// ---------------------------------------------------------------------------
// RK4-Nystroem step on the free vector F = [x, y, z, t, Tx, Ty, Tz, q/p] and
// its 8x8 transport matrix D = dF'/dF (field gradient neglected).
// ---------------------------------------------------------------------------
template <typename S>
S dt_ds(S qop, S mass, S abs_q) {
    const S mq = mass * qop / abs_q;
    return std::sqrt(S(1) + mq * mq) / static_cast<S>(kC);
}

template <typename S>
void rk_step(const ckf_bfield& field, S mass, S abs_q, S h, mat<S, 8, 1>& F, mat<S, 8, 8>* D) {
    const v3<S> r0 = mk3<S>(F.a[0][0], F.a[1][0], F.a[2][0]);
    const v3<S> T0 = mk3<S>(F.a[4][0], F.a[5][0], F.a[6][0]);
    const S q = F.a[7][0];
    const S half = h * S(0.5);

    const v3<S> B1 = bfield<S>(field, r0);
    const v3<S> k1 = scale3(q, cross3(T0, B1));
    const v3<S> r2 = r0 + scale3(half, T0) + scale3(h * h * S(0.125), k1);
    const v3<S> B2 = bfield<S>(field, r2);
    const v3<S> T2 = T0 + scale3(half, k1);
    const v3<S> k2 = scale3(q, cross3(T2, B2));
    const v3<S> T3 = T0 + scale3(half, k2);
    const v3<S> k3 = scale3(q, cross3(T3, B2));
    const v3<S> r4 = r0 + scale3(h, T0) + scale3(h * h * S(0.5), k3);
    const v3<S> B4 = bfield<S>(field, r4);
    const v3<S> T4 = T0 + scale3(h, k3);
    const v3<S> k4 = scale3(q, cross3(T4, B4));

    if (D != nullptr) {
        // dk_i/dF as 3x8 blocks.  k = q T_i x B_i; dT_i/dF from the previous stage.
        auto xB = [&](const mat<S, 3, 8>& dT, const v3<S>& B) {
            // d(q * T x B)/dF restricted to the T dependence
            mat<S, 3, 8> m;
            for (int j = 0; j < 8; ++j) {
                const v3<S> col = mk3<S>(dT.a[0][j], dT.a[1][j], dT.a[2][j]);
                const v3<S> c = scale3(q, cross3(col, B));
                m.a[0][j] = c.a[0][0];
                m.a[1][j] = c.a[1][0];
                m.a[2][j] = c.a[2][0];
            }
            return m;
        };
        mat<S, 3, 8> dT1 = mat<S, 3, 8>::zero();
        dT1.a[0][4] = dT1.a[1][5] = dT1.a[2][6] = S(1);
        mat<S, 3, 8> dk1 = xB(dT1, B1);
        {
            const v3<S> c = cross3(T0, B1);
            for (int i = 0; i < 3; ++i) dk1.a[i][7] += c.a[i][0];
        }
        mat<S, 3, 8> dT2 = dT1;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 8; ++j) dT2.a[i][j] += half * dk1.a[i][j];
        mat<S, 3, 8> dk2 = xB(dT2, B2);
        {
            const v3<S> c = cross3(T2, B2);
            for (int i = 0; i < 3; ++i) dk2.a[i][7] += c.a[i][0];
        }
        mat<S, 3, 8> dT3 = dT1;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 8; ++j) dT3.a[i][j] += half * dk2.a[i][j];
        mat<S, 3, 8> dk3 = xB(dT3, B2);
        {
            const v3<S> c = cross3(T3, B2);
            for (int i = 0; i < 3; ++i) dk3.a[i][7] += c.a[i][0];
        }
        mat<S, 3, 8> dT4 = dT1;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 8; ++j) dT4.a[i][j] += h * dk3.a[i][j];
        mat<S, 3, 8> dk4 = xB(dT4, B4);
        {
            const v3<S> c = cross3(T4, B4);
            for (int i = 0; i < 3; ++i) dk4.a[i][7] += c.a[i][0];
        }
        mat<S, 8, 8> M = mat<S, 8, 8>::identity();
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 8; ++j) {
                // position rows: r' = r + h T + h^2/6 (k1 + k2 + k3)
                M.a[i][j] += (j == 4 + i ? h : S(0)) +
                             h * h / S(6) * (dk1.a[i][j] + dk2.a[i][j] + dk3.a[i][j]);
                // direction rows: T' = T + h/6 (k1 + 2 k2 + 2 k3 + k4)
                M.a[4 + i][j] += h / S(6) *
                                 (dk1.a[i][j] + S(2) * dk2.a[i][j] + S(2) * dk3.a[i][j] + dk4.a[i][j]);
            }
        }
        // time row: t' = t + h / (beta c)
        const S mq = mass / abs_q;
        M.a[3][7] += h * mq * mq * q / (static_cast<S>(kC) * static_cast<S>(kC) * dt_ds(q, mass, abs_q));
        *D = M * (*D);
    }

    const v3<S> rn = r0 + scale3(h, T0) + scale3(h * h / S(6), k1 + k2 + k3);
    v3<S> Tn = T0 + scale3(h / S(6), k1 + scale3(S(2), k2) + scale3(S(2), k3) + k4);
    Tn = scale3(S(1) / std::sqrt(dot3(Tn, Tn)), Tn);
    F.a[0][0] = rn.a[0][0];
    F.a[1][0] = rn.a[1][0];
    F.a[2][0] = rn.a[2][0];
    F.a[3][0] += h * dt_ds(q, mass, abs_q);
    F.a[4][0] = Tn.a[0][0];
    F.a[5][0] = Tn.a[1][0];
    F.a[6][0] = Tn.a[2][0];
}

// ---------------------------------------------------------------------------
// Geometry queries
// ---------------------------------------------------------------------------
template <typename S>
bool in_mask(const ckf_surface& s, S l0, S l1, S tol) {
    if (s.bound_kind == CKF_BOUND_RECTANGLE) {
        return std::fabs(l0) <= static_cast<S>(s.bound[0]) + tol &&
               std::fabs(l1) <= static_cast<S>(s.bound[1]) + tol;
    }
    const S hh = static_cast<S>(s.bound[2]);
    if (std::fabs(l1) > hh + tol) return false;
    const S lim = static_cast<S>(s.bound[0]) +
                  (static_cast<S>(s.bound[1]) - static_cast<S>(s.bound[0])) * (l1 + hh) / (S(2) * hh);
    return std::fabs(l0) <= lim + tol;
}

// Straight-line entry distance into a layer envelope (0 inside, +inf never).
template <typename S>
S envelope_entry(const ckf_layer& L, const v3<S>& p, const v3<S>& d, S margin) {
    const S inf = static_cast<S>(1e30);
    if (L.kind == CKF_LAYER_PLANE) {
        const v3<S> n = fl3<S>(L.normal);
        const S lo = static_cast<S>(L.env_r[0]), hi = static_cast<S>(L.env_r[1]);
        const S here = dot3(n, p);
        if (here >= lo && here <= hi) return S(0);
        const S rate = dot3(n, d);
        if (std::fabs(rate) < static_cast<S>(1e-12)) return inf;
        S best = inf;
        for (S face : {lo, hi}) {
            const S s = (face - here) / rate;
            if (s > S(0) && s < best) best = s;
        }
        return best;
    }
    const S rlo = static_cast<S>(L.env_r[0]), rhi = static_cast<S>(L.env_r[1]);
    const S zlo = static_cast<S>(L.env_z[0]), zhi = static_cast<S>(L.env_z[1]);
    const S px = p.a[0][0], py = p.a[1][0], pz = p.a[2][0];
    const S dx = d.a[0][0], dy = d.a[1][0], dz = d.a[2][0];
    const S rho2 = px * px + py * py;
    if (rho2 >= rlo * rlo && rho2 <= rhi * rhi && pz >= zlo && pz <= zhi) return S(0);
    S best = inf;
    const S A = dx * dx + dy * dy;
    if (A > static_cast<S>(1e-12)) {
        const S Bh = px * dx + py * dy;
        for (S rad : {rlo, rhi}) {
            const S disc = Bh * Bh - A * (rho2 - rad * rad);
            if (disc < S(0)) continue;
            const S sq = std::sqrt(disc);
            for (S s : {(-Bh - sq) / A, (-Bh + sq) / A}) {
                if (s > S(0) && s < best) {
                    const S zc = pz + s * dz;
                    if (zc >= zlo - margin && zc <= zhi + margin) best = s;
                }
            }
        }
    }
    if (std::fabs(dz) > static_cast<S>(1e-12)) {
        for (S face : {zlo, zhi}) {
            const S s = (face - pz) / dz;
            if (s > S(0) && s < best) {
                const S xc = px + s * dx, yc = py + s * dy;
                const S rr = xc * xc + yc * yc;
                const S lo = std::max(rlo - margin, S(0)), hi = rhi + margin;
                if (rr >= lo * lo && rr <= hi * hi) best = s;
            }
        }
    }
    return best;
}

// ---------------------------------------------------------------------------
// Propagation engine for one track
// ---------------------------------------------------------------------------
template <typename S>
struct propagation {
    const ckf_finding_config& cfg;
    const ckf_detector& det;
    const ckf_bfield& field;
    bool with_jacobian;
    S mass, abs_q;

    mat<S, 8, 1> F;
    mat<S, 8, 8> D;  // accumulated transport
    S path = S(0);
    uint32_t steps = 0;
    // a surface is reached when the crossing lies within this distance of its
    // mask (0: the exact mask, the CKF; the fit: min_mask_tolerance)
    S accept_tol = S(0);

    S step_scale() const {
        return static_cast<S>(std::pow(120.0 * static_cast<double>(cfg.propagation.rk_error_tol), 0.2));
    }

    v3<S> pos() const { return mk3<S>(F.a[0][0], F.a[1][0], F.a[2][0]); }
    v3<S> dir() const { return mk3<S>(F.a[4][0], F.a[5][0], F.a[6][0]); }

    S accuracy_step() const {
        const v3<S> B = bfield<S>(field, pos());
        const v3<S> tb = cross3(dir(), B);
        const S kappa = std::fabs(F.a[7][0]) * std::sqrt(dot3(tb, tb));
        S h = static_cast<S>(cfg.propagation.step_constraint);
        if (kappa > static_cast<S>(1e-12)) {
            h = std::min(h, step_scale() * std::pow(kappa, static_cast<S>(-0.8)));
        }
        return h;
    }

    bool step(S h) {
        rk_step<S>(field, mass, abs_q, h, F, with_jacobian ? &D : nullptr);
        path += h;
        ++steps;
        return steps <= cfg.max_step_counts_for_next_surface &&
               std::fabs(path) <= static_cast<S>(cfg.propagation.path_limit);
    }

    // 0 on plane, 1 abort, 2 parallel
    int onto_plane(const ckf_surface& s) {
        const v3<S> c = fl3<S>(s.center);
        const v3<S> w = fl3<S>(s.normal);
        S dist = S(0);
        while (true) {
            const S cosw = dot3(dir(), w);
            if (std::fabs(cosw) < static_cast<S>(1e-8)) return 2;
            dist = dot3(c - pos(), w) / cosw;
            if (std::fabs(dist) <= static_cast<S>(cfg.propagation.on_surface_tolerance)) break;
            const S lim = accuracy_step();
            const S h = dist > S(0) ? std::min(dist, lim) : std::max(dist, -lim);
            if (!step(h)) return 1;
        }
        // the remaining sub-tolerance distance is one more RK step
        if (dist != S(0) && !step(dist)) return 1;
        return 0;
    }

    // 0 found, 1 none, 2 abort
    int scan_layer(uint32_t li, int skip, uint32_t& found) {
        const ckf_layer& L = det.layers[li];
        const S margin = static_cast<S>(cfg.propagation.max_mask_tolerance);
        S se = envelope_entry<S>(L, pos(), dir(), margin);
        if (!(se < static_cast<S>(1e29))) se = S(0);
        const v3<S> at = pos() + scale3(se, dir());
        const int nphi = static_cast<int>(L.n_phi), nlong = static_cast<int>(L.n_long);
        struct cand {
            S s;
            uint32_t idx;
        };
        cand plane_cands[9];
        size_t n_plane = 0;
        int cphi = 0, clong = 0;
        int wphi = static_cast<int>(cfg.propagation.search_window[0]);
        int wlong = L.kind == CKF_LAYER_BARREL ? static_cast<int>(cfg.propagation.search_window[1]) : 0;
        if (L.kind == CKF_LAYER_PLANE) {
            // every module of a plane layer (up to 9) is a candidate
            for (int k = 0; k < nphi * nlong && k < 9; ++k) {
                const uint32_t m = L.first_surface + static_cast<uint32_t>(k);
                if (static_cast<int>(m) == skip) continue;
                const ckf_surface& sf = det.surfaces[m];
                const v3<S> w = fl3<S>(sf.normal);
                const S cosw = dot3(dir(), w);
                if (std::fabs(cosw) < static_cast<S>(1e-6)) continue;
                const S s = dot3(fl3<S>(sf.center) - pos(), w) / cosw;
                if (s < static_cast<S>(cfg.propagation.overstep_tolerance)) continue;
                const v3<S> rel = pos() + scale3(s, dir()) - fl3<S>(sf.center);
                const S tol = std::min(static_cast<S>(cfg.propagation.min_mask_tolerance) +
                                           static_cast<S>(cfg.propagation.mask_tolerance_scalor) * std::fabs(s),
                                       static_cast<S>(cfg.propagation.max_mask_tolerance));
                if (!in_mask<S>(sf, dot3(rel, fl3<S>(sf.axis_u)), dot3(rel, fl3<S>(sf.axis_v)), tol)) continue;
                plane_cands[n_plane++] = {s, m};
            }
            wphi = -1;  // no window scan below
        } else {
            const S ph = std::atan2(at.a[1][0], at.a[0][0]);
            const S pitch = static_cast<S>(2.0 * kPi) / static_cast<S>(nphi);
            cphi = static_cast<int>(std::floor((ph - static_cast<S>(L.phi0)) / pitch + S(0.5))) % nphi;
            if (cphi < 0) cphi += nphi;
            if (L.kind == CKF_LAYER_BARREL && nlong > 1) {
                clong = static_cast<int>(std::floor((at.a[2][0] - static_cast<S>(L.long0)) /
                                                        static_cast<S>(L.long_pitch) + S(0.5)));
                clong = std::min(std::max(clong, 0), nlong - 1);
            }
        }
        if (wphi >= 0 && 2 * wphi + 1 > nphi) wphi = (nphi - 1) / 2;
        cand cbuf[9];
        size_t nc = 0;
        for (size_t k = 0; k < n_plane; ++k) cbuf[nc++] = plane_cands[k];
        for (int a = -wphi; a <= wphi; ++a) {
            const int ip = ((cphi + a) % nphi + nphi) % nphi;
            for (int b = -wlong; b <= wlong; ++b) {
                const int il = clong + b;
                if (il < 0 || il >= nlong) continue;
                const uint32_t m = L.first_surface + static_cast<uint32_t>(ip * nlong + il);
                if (static_cast<int>(m) == skip) continue;
                const ckf_surface& sf = det.surfaces[m];
                const v3<S> w = fl3<S>(sf.normal);
                const S cosw = dot3(dir(), w);
                if (std::fabs(cosw) < static_cast<S>(1e-6)) continue;
                const S s = dot3(fl3<S>(sf.center) - pos(), w) / cosw;
                if (s < static_cast<S>(cfg.propagation.overstep_tolerance)) continue;
                const v3<S> rel = pos() + scale3(s, dir()) - fl3<S>(sf.center);
                const S tol = std::min(static_cast<S>(cfg.propagation.min_mask_tolerance) +
                                           static_cast<S>(cfg.propagation.mask_tolerance_scalor) * std::fabs(s),
                                       static_cast<S>(cfg.propagation.max_mask_tolerance));
                if (!in_mask<S>(sf, dot3(rel, fl3<S>(sf.axis_u)), dot3(rel, fl3<S>(sf.axis_v)), tol)) continue;
                if (nc < 9u) cbuf[nc++] = {s, m};
            }
        }
        std::sort(cbuf, cbuf + nc, [](const cand& x, const cand& y) {
            return x.s < y.s || (x.s == y.s && x.idx < y.idx);
        });
        for (size_t ci = 0; ci < nc; ++ci) {
            const cand& c = cbuf[ci];
            const ckf_surface& sf = det.surfaces[c.idx];
            const v3<S> w = fl3<S>(sf.normal);
            const S cosw = dot3(dir(), w);
            if (std::fabs(cosw) < static_cast<S>(1e-8)) continue;
            if (dot3(fl3<S>(sf.center) - pos(), w) / cosw < static_cast<S>(cfg.propagation.overstep_tolerance)) continue;
            const int r = onto_plane(sf);
            if (r == 1) return 2;
            if (r == 2) continue;
            const v3<S> rel = pos() - fl3<S>(sf.center);
            if (in_mask<S>(sf, dot3(rel, fl3<S>(sf.axis_u)), dot3(rel, fl3<S>(sf.axis_v)), accept_tol) &&
                path > static_cast<S>(cfg.min_step_length_for_next_surface)) {
                found = c.idx;
                return 0;
            }
        }
        return 1;
    }

    // 0 found, 1 exit, 2 abort
    int run(int start, uint32_t& found) {
        path = S(0);
        steps = 0u;
        const S p = abs_q / std::fabs(F.a[7][0]);
        const S pT = p * std::sqrt(F.a[4][0] * F.a[4][0] + F.a[5][0] * F.a[5][0]);
        if (p < static_cast<S>(cfg.min_p) || pT < static_cast<S>(cfg.min_pT)) return 2;
        uint64_t done = 0ull;
        if (start >= 0) {
            const uint32_t l0 = det.surfaces[start].layer;
            const int r = scan_layer(l0, start, found);
            if (r == 0) return 0;
            if (r == 2) return 2;
            done |= 1ull << l0;
        }
        const S margin = static_cast<S>(cfg.propagation.max_mask_tolerance);
        const S approach = S(25);
        for (uint32_t round = 0; round <= det.n_layers; ++round) {
            int next = -1;
            S dist = static_cast<S>(1e30);
            for (uint32_t li = 0; li < det.n_layers; ++li) {
                if (done & (1ull << li)) continue;
                const S s = envelope_entry<S>(det.layers[li], pos(), dir(), margin);
                if (s < dist) {
                    dist = s;
                    next = static_cast<int>(li);
                }
            }
            if (next < 0 || !(dist < static_cast<S>(1e29))) return 1;
            while (dist > approach && dist < static_cast<S>(1e29)) {
                if (!step(std::min(dist, accuracy_step()))) return 2;
                dist = envelope_entry<S>(det.layers[next], pos(), dir(), margin);
            }
            if (dist < static_cast<S>(1e29)) {
                const int r = scan_layer(static_cast<uint32_t>(next), -1, found);
                if (r == 0) return 0;
                if (r == 2) return 2;
            }
            done |= 1ull << next;
        }
        return 1;
    }
};

template <typename S>
mat<S, 8, 6> bound_to_free_jac(const ckf_surface& sf, const mat<S, 6, 1>& v) {
    const S phi = v.a[2][0], th = v.a[3][0];
    mat<S, 8, 6> J = mat<S, 8, 6>::zero();
    for (int i = 0; i < 3; ++i) {
        J.a[i][0] = static_cast<S>(sf.axis_u[i]);
        J.a[i][1] = static_cast<S>(sf.axis_v[i]);
    }
    J.a[3][5] = S(1);
    J.a[4][2] = -std::sin(phi) * std::sin(th);
    J.a[5][2] = std::cos(phi) * std::sin(th);
    J.a[6][2] = S(0);
    J.a[4][3] = std::cos(phi) * std::cos(th);
    J.a[5][3] = std::sin(phi) * std::cos(th);
    J.a[6][3] = -std::sin(th);
    J.a[7][4] = S(1);
    return J;
}

// Highland width (PDG eq. 34.15)
template <typename S>
S highland(S x, S qop, S mass, S abs_q) {
    if (!(x > S(0))) return S(0);
    const S p = abs_q / std::fabs(qop);
    const S beta = p / std::sqrt(p * p + mass * mass);
    return S(0.0136) / (beta * p) * abs_q * std::sqrt(x) *
           (S(1) + S(0.038) * std::log(x * abs_q * abs_q / (beta * beta)));
}

// Multiple-scattering noise of the module the state sits on
// (pointwise material interaction, covariance part).
template <typename S>
void add_material(const ckf_surface& sf, bound_state<S>& st, S mass, S abs_q) {
    if (!(sf.thickness_x0 > 0.f)) return;
    const S phi = st.vec.a[2][0], th = st.vec.a[3][0];
    const v3<S> d = mk3<S>(std::cos(phi) * std::sin(th), std::sin(phi) * std::sin(th), std::cos(th));
    const S cosi = std::fabs(dot3(d, fl3<S>(sf.normal)));
    const S t0 = highland<S>(static_cast<S>(sf.thickness_x0) / cosi, st.vec.a[4][0], mass, abs_q);
    const S sth = std::sin(th);
    st.cov.a[2][2] += t0 * t0 / (sth * sth);
    st.cov.a[3][3] += t0 * t0;
}

// Free state of a propagation engine from bound parameters on surface s0.
template <typename S>
void load_free(propagation<S>& P, const ckf_surface& s0, const bound_state<S>& in) {
    const S l0 = in.vec.a[0][0], l1 = in.vec.a[1][0], phi = in.vec.a[2][0], th = in.vec.a[3][0];
    for (int i = 0; i < 3; ++i) {
        P.F.a[i][0] = static_cast<S>(s0.center[i]) + l0 * static_cast<S>(s0.axis_u[i]) +
                      l1 * static_cast<S>(s0.axis_v[i]);
    }
    P.F.a[3][0] = in.vec.a[5][0];
    P.F.a[4][0] = std::cos(phi) * std::sin(th);
    P.F.a[5][0] = std::sin(phi) * std::sin(th);
    P.F.a[6][0] = std::cos(th);
    P.F.a[7][0] = in.vec.a[4][0];
    P.D = mat<S, 8, 8>::identity();
}

// Bound parameters on surface s1 (index `found`) from the engine's free state,
// with the covariance carried through d(bound_1)/d(bound_0).
template <typename S>
void finish_bound(const propagation<S>& P, const ckf_bfield& field, const ckf_surface& s0,
                  const ckf_surface& s1, uint32_t found, const bound_state<S>& in, S mass, S abs_q,
                  bound_state<S>& out) {
    const v3<S> rel = P.pos() - fl3<S>(s1.center);
    const v3<S> T = P.dir();
    out.surface = found;
    out.vec.a[0][0] = dot3(rel, fl3<S>(s1.axis_u));
    out.vec.a[1][0] = dot3(rel, fl3<S>(s1.axis_v));
    out.vec.a[2][0] = std::atan2(T.a[1][0], T.a[0][0]);
    out.vec.a[3][0] = std::acos(std::min(std::max(T.a[2][0], S(-1)), S(1)));
    out.vec.a[4][0] = P.F.a[7][0];
    out.vec.a[5][0] = P.F.a[3][0];

    // path correction onto the plane: dF -> dF - (dF/ds) (w . dr)/(w . T)
    const v3<S> w = fl3<S>(s1.normal);
    const S wt = dot3(w, T);
    const v3<S> B = bfield<S>(field, P.pos());
    const v3<S> dTds = scale3(P.F.a[7][0], cross3(T, B));
    mat<S, 8, 1> dFds = mat<S, 8, 1>::zero();
    for (int i = 0; i < 3; ++i) {
        dFds.a[i][0] = T.a[i][0];
        dFds.a[4 + i][0] = dTds.a[i][0];
    }
    dFds.a[3][0] = dt_ds(P.F.a[7][0], mass, abs_q);
    mat<S, 1, 8> dsdF = mat<S, 1, 8>::zero();
    for (int i = 0; i < 3; ++i) dsdF.a[0][i] = -w.a[i][0] / wt;
    const mat<S, 8, 8> corr = mat<S, 8, 8>::identity() + dFds * dsdF;

    // free -> bound projection
    mat<S, 6, 8> Jfb = mat<S, 6, 8>::zero();
    const S sin2 = T.a[0][0] * T.a[0][0] + T.a[1][0] * T.a[1][0];
    const S sinT = std::sqrt(sin2);
    for (int i = 0; i < 3; ++i) {
        Jfb.a[0][i] = static_cast<S>(s1.axis_u[i]);
        Jfb.a[1][i] = static_cast<S>(s1.axis_v[i]);
    }
    Jfb.a[2][4] = -T.a[1][0] / sin2;
    Jfb.a[2][5] = T.a[0][0] / sin2;
    Jfb.a[3][4] = T.a[0][0] * T.a[2][0] / sinT;
    Jfb.a[3][5] = T.a[1][0] * T.a[2][0] / sinT;
    Jfb.a[3][6] = -sinT;
    Jfb.a[4][7] = S(1);
    Jfb.a[5][3] = S(1);

    const mat<S, 6, 6> J = Jfb * corr * P.D * bound_to_free_jac<S>(s0, in.vec);
    out.cov = J * in.cov * transpose(J);
}

// Propagate a bound state to the next sensitive surface.
// 0 found (out filled), 1 exit, 2 abort.
template <typename S>
int propagate_bound(const ckf_finding_config& cfg, const ckf_detector& det, const ckf_bfield& field,
                    const bound_state<S>& in, S mass, S abs_q, bound_state<S>& out, S* path_out,
                    uint32_t* steps_out, S accept_tol = S(0)) {
    const ckf_surface& s0 = det.surfaces[in.surface];
    propagation<S> P{cfg, det, field, true, mass, abs_q, mat<S, 8, 1>::zero(), mat<S, 8, 8>::identity(),
                     S(0), 0u};
    P.accept_tol = accept_tol;
    load_free<S>(P, s0, in);
    uint32_t found = 0u;
    const int r = P.run(static_cast<int>(in.surface), found);
    if (path_out) *path_out = P.path;
    if (steps_out) *steps_out = P.steps;
    if (r != 0) return r;
    finish_bound<S>(P, field, s0, det.surfaces[found], found, in, mass, abs_q, out);
    return 0;
}

// Propagate a bound state straight onto surface `target`, forwards or
// backwards along the track (the direct navigation of a fit's surface
// sequence).  0 on the surface (out filled), 2 abort.
template <typename S>
int propagate_direct(const ckf_finding_config& cfg, const ckf_detector& det, const ckf_bfield& field,
                     const bound_state<S>& in, uint32_t target, S mass, S abs_q, bound_state<S>& out) {
    const ckf_surface& s0 = det.surfaces[in.surface];
    propagation<S> P{cfg, det, field, true, mass, abs_q, mat<S, 8, 1>::zero(), mat<S, 8, 8>::identity(),
                     S(0), 0u};
    load_free<S>(P, s0, in);
    if (P.onto_plane(det.surfaces[target]) != 0) return 2;
    finish_bound<S>(P, field, s0, det.surfaces[target], target, in, mass, abs_q, out);
    return 0;
}

// ---------------------------------------------------------------------------
// chi2 p-value: upper regularised incomplete gamma Q(ndf/2, chi2/2), the
// series / continued-fraction split of the reference's Cephes-derived code
// (core/include/traccc/utils/prob.hpp:148-288), log-gamma from the C library.
// ---------------------------------------------------------------------------
template <typename S>
S igam_series(S a, S x) {
    S ax = a * std::log(x) - x - static_cast<S>(std::lgamma(static_cast<double>(a)));
    if (ax < static_cast<S>(-709.78)) return S(0);
    ax = std::exp(ax);
    S r = a, c = S(1), ans = S(1);
    do {
        r += S(1);
        c *= x / r;
        ans += c;
    } while (c / ans > static_cast<S>(1.11022302462515654042e-16));
    return ans * ax / a;
}

template <typename S>
S igamc_fraction(S a, S x) {
    S ax = a * std::log(x) - x - static_cast<S>(std::lgamma(static_cast<double>(a)));
    if (ax < static_cast<S>(-709.78)) return S(0);
    ax = std::exp(ax);
    S y = S(1) - a, z = x + y + S(1), c = S(0);
    S pkm2 = S(1), qkm2 = x, pkm1 = x + S(1), qkm1 = z * x;
    S ans = pkm1 / qkm1, t;
    const S big = static_cast<S>(4.503599627370496e15);
    const S biginv = static_cast<S>(2.22044604925031308085e-16);
    int guard = 0;
    do {
        c += S(1);
        y += S(1);
        z += S(2);
        const S yc = y * c;
        const S pk = pkm1 * z - pkm2 * yc;
        const S qk = qkm1 * z - qkm2 * yc;
        if (qk != S(0)) {
            const S r = pk / qk;
            t = std::fabs((ans - r) / r);
            ans = r;
        } else {
            t = S(1);
        }
        pkm2 = pkm1;
        pkm1 = pk;
        qkm2 = qkm1;
        qkm1 = qk;
        if (std::fabs(pk) > big) {
            pkm2 *= biginv;
            pkm1 *= biginv;
            qkm2 *= biginv;
            qkm1 *= biginv;
        }
    } while (t > static_cast<S>(1.11022302462515654042e-16) && ++guard < 2000);
    return ans * ax;
}

template <typename S>
S prob(S chi2, S ndf) {
    if (ndf <= S(0)) return S(0);
    if (chi2 <= S(0)) return chi2 < S(0) ? S(0) : S(1);
    const S a = S(0.5) * ndf, x = S(0.5) * chi2;
    if (x < S(1) || x < a) return S(1) - igam_series<S>(a, x);
    return igamc_fraction<S>(a, x);
}

// ---------------------------------------------------------------------------
// Kalman filter pieces (measurement_selector.hpp:53-212, gain_matrix_updater.hpp:69-181)
// ---------------------------------------------------------------------------
template <typename S>
struct meas_view {
    S pos[2];  // already reordered by the subspace (get_measurement_local)
    S var[2];
    uint32_t dims;
    uint8_t subs[2];
};

template <typename S>
meas_view<S> load_meas(const ckf_measurements& m, uint32_t i) {
    meas_view<S> v;
    v.dims = m.dimensions[i];
    v.subs[0] = m.subspace[2 * i];
    v.subs[1] = m.subspace[2 * i + 1];
    const S a = static_cast<S>(m.local_position[2 * i]), b = static_cast<S>(m.local_position[2 * i + 1]);
    const S va = static_cast<S>(m.local_variance[2 * i]), vb = static_cast<S>(m.local_variance[2 * i + 1]);
    // measurement_helpers.ipp: subspace[0] == loc1 swaps the two coordinates
    if (v.subs[0] == 1u) {
        v.pos[0] = b;
        v.pos[1] = a;
        v.var[0] = vb;
        v.var[1] = va;
    } else {
        v.pos[0] = a;
        v.pos[1] = b;
        v.var[0] = va;
        v.var[1] = vb;
    }
    if (v.dims == 1u) v.var[1] = static_cast<S>(std::numeric_limits<float>::max());
    return v;
}

template <typename S>
mat<S, 2, 6> projector(const meas_view<S>& m) {
    mat<S, 2, 6> H = mat<S, 2, 6>::zero();
    H.a[0][m.subs[0]] = S(1);
    if (m.dims == 2u) H.a[1][m.subs[1]] = S(1);
    return H;
}

template <typename S>
mat<S, 2, 2> masked_inv(const mat<S, 2, 2>& M, uint32_t dims) {
    mat<S, 2, 2> R = mat<S, 2, 2>::zero();
    if (dims == 1u) {
        R.a[0][0] = S(1) / M.a[0][0];
        return R;
    }
    const S det = M.a[0][0] * M.a[1][1] - M.a[0][1] * M.a[1][0];
    R.a[0][0] = M.a[1][1] / det;
    R.a[1][1] = M.a[0][0] / det;
    R.a[0][1] = -M.a[0][1] / det;
    R.a[1][0] = -M.a[1][0] / det;
    return R;
}

template <typename S>
S predicted_chi2(const meas_view<S>& m, const bound_state<S>& p) {
    const mat<S, 2, 6> H = projector(m);
    mat<S, 2, 2> V = mat<S, 2, 2>::zero();
    V.a[0][0] = m.var[0];
    V.a[1][1] = m.var[1];
    const mat<S, 2, 2> R = H * (p.cov * transpose(H)) + V;
    const mat<S, 2, 2> Ri = masked_inv(R, m.dims);
    mat<S, 2, 1> ml;
    ml.a[0][0] = m.pos[0];
    ml.a[1][0] = m.dims == 2u ? m.pos[1] : S(0);
    const mat<S, 2, 1> res = ml - H * p.vec;
    return (transpose(res) * Ri * res).a[0][0];
}

template <typename S>
bool wrap_angles(mat<S, 6, 1>& v) {
    const S twopi = static_cast<S>(2.0 * kPi), pi = static_cast<S>(kPi);
    const S th_in = v.a[3][0];
    if (std::fmod(th_in, twopi) == S(0) || th_in == S(0)) return false;
    S phi = v.a[2][0];
    S th = th_in;
    // wrap_periodic(theta, 0, 2pi)
    if (!(th >= S(0) && th < twopi)) th = th - twopi * std::floor(th / twopi);
    if (pi < th) {
        phi += pi;
        th = twopi - th;
    }
    const S d = phi + pi;
    if (!(d >= S(0) && d < twopi)) phi = phi - twopi * std::floor(d / twopi);
    if (th <= S(0) || th >= twopi) return false;
    v.a[2][0] = phi;
    v.a[3][0] = th;
    return true;
}

template <typename S>
int gain_update(const meas_view<S>& m, const bound_state<S>& pred, bound_state<S>& filt) {
    const mat<S, 2, 6> H = projector(m);
    mat<S, 2, 2> V = mat<S, 2, 2>::zero();
    V.a[0][0] = m.var[0];
    V.a[1][1] = m.var[1];
    const mat<S, 6, 2> PHt = pred.cov * transpose(H);
    const mat<S, 2, 2> Minv = masked_inv(H * PHt + V, m.dims);
    const mat<S, 6, 2> K = PHt * Minv;
    mat<S, 2, 1> ml;
    ml.a[0][0] = m.pos[0];
    ml.a[1][0] = m.dims == 2u ? m.pos[1] : S(0);
    mat<S, 6, 1> x = pred.vec + K * (ml - H * pred.vec);
    if (!std::isfinite(x.a[3][0]) || !std::isfinite(x.a[2][0])) return CKF_FIT_ERROR_INVERSION;
    if (std::fabs(x.a[4][0]) == S(0)) return CKF_FIT_ERROR_QOP_ZERO;
    if (!wrap_angles(x)) return CKF_FIT_ERROR_THETA_POLE;
    const mat<S, 6, 6> IKH = mat<S, 6, 6>::identity() - K * H;
    mat<S, 6, 6> C = IKH * pred.cov * transpose(IKH) + K * V * transpose(K);
    const S min_var = S(-0.01);
    for (int i = 0; i < 6; ++i)
        if (C.a[i][i] < min_var) return CKF_FIT_ERROR_UPDATER_INVALID_COVARIANCE;
    for (int i = 0; i < 6; ++i) C.a[i][i] = std::fabs(C.a[i][i]);
    filt.surface = pred.surface;
    filt.vec = x;
    filt.cov = C;
    return CKF_FIT_SUCCESS;
}

// ---------------------------------------------------------------------------
// The CKF (host semantics)
// ---------------------------------------------------------------------------
struct link {
    uint32_t step;
    uint32_t previous_candidate_idx;
    uint32_t meas_idx;
    uint32_t seed_idx;
    uint32_t n_skipped;
    uint32_t n_consecutive_skipped;
    double chi2;
    double chi2_sum;
    uint32_t ndf_sum;
};

struct track_out {
    uint32_t seed;
    double ndf, chi2, pval;
    double vote_pval;  // p-value of the tip link's running sums (measurement votes)
    uint32_t nholes;
    std::vector<uint32_t> meas;
};

struct run_stats {
    uint32_t steps = 0;
    uint64_t links = 0, props = 0, tips = 0;
};

// Candidate verdicts of a pruned run (the learned branch scorer's decisions,
// taken from the device): per input only the list_len best candidates by
// (chi2, index) that pass the cut are listed, and of those only the ones whose
// key is in `kept` (sorted) go to the gain update, in that order, until
// max_num_branches_per_surface succeed.  A key names a candidate by its
// branch's history: splitmix64 (Steele, Lea, Flood 2014) chained as
// h = mix(h ^ v) over the seed, the step, the measurement of every earlier
// link of the branch (latest first, holes as all-ones) and the candidate's
// measurement, low 63 bits.
inline uint64_t verdict_mix(uint64_t h, uint64_t v) {
    uint64_t z = (h ^ v) + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct cand_mask {
    const uint64_t* kept = nullptr;
    size_t n_kept = 0;
    uint32_t list_len = 0;
    bool keep_all = false;
    bool has(uint64_t branch, uint32_t m) const {
        if (keep_all) return true;
        const uint64_t k = verdict_mix(branch, m) & 0x7fffffffffffffffull;
        return std::binary_search(kept, kept + n_kept, k);
    }
};

// surface identifier -> index
inline uint32_t surface_index(const ckf_detector& det, uint64_t id) {
    uint32_t lo = 0, hi = det.n_surfaces;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (det.surfaces[mid].identifier < id) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename S>
std::vector<track_out> ckf(const ckf_finding_config& cfg, const ckf_detector& det,
                           const ckf_bfield& field, const ckf_measurements& meas,
                           const ckf_bound_params& seeds, run_stats& stats,
                           const cand_mask* mask = nullptr) {
    const uint32_t n_meas = meas.size;
    const S mass = static_cast<S>(cfg.ptc_mass);
    const S abs_q = static_cast<S>(std::fabs(cfg.ptc_charge));
    const S chi2_max = static_cast<S>(cfg.chi2_max);

    // measurement ranges per surface (upper bounds), reference :97-120
    std::vector<uint32_t> meas_ranges(det.n_surfaces);
    for (uint32_t s = 0; s < det.n_surfaces; ++s) {
        if (!det.surfaces[s].sensitive) {
            meas_ranges[s] = s == 0 ? 0u : meas_ranges[s - 1];
            continue;
        }
        const uint64_t id = det.surfaces[s].identifier;
        meas_ranges[s] = static_cast<uint32_t>(
            std::upper_bound(meas.surface_link, meas.surface_link + n_meas, id) - meas.surface_link);
    }

    const uint32_t n_steps_max = cfg.max_track_candidates_per_track;
    std::vector<std::vector<link>> links(n_steps_max);
    std::vector<std::vector<size_t>> param_to_link(n_steps_max);
    std::vector<std::pair<uint32_t, uint32_t>> tips;

    std::vector<bound_state<S>> in_params(seeds.size);
    for (uint32_t i = 0; i < seeds.size; ++i) {
        in_params[i].surface = surface_index(det, seeds.surface_link[i]);
        for (int k = 0; k < 6; ++k) in_params[i].vec.a[k][0] = static_cast<S>(seeds.vector[6 * i + k]);
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) in_params[i].cov.a[r][c] = static_cast<S>(seeds.covariance[36 * i + 6 * r + c]);
    }
    std::vector<uint32_t> n_trks_per_seed(seeds.size);
    std::vector<bound_state<S>> out_params;
    // the link a link continues (param_to_link maps an input index to its link)
    auto back = [&](const link& L) {
        const size_t pos = param_to_link[L.step - 1][L.previous_candidate_idx];
        return links[L.step - 1][pos];
    };

    for (uint32_t step = 0; step < n_steps_max; ++step) {
        const size_t n_in = in_params.size();
        if (n_in == 0) break;
        stats.steps = step + 1;
        std::fill(n_trks_per_seed.begin(), n_trks_per_seed.end(), 0u);
        std::vector<bound_state<S>> updated;

        for (uint32_t ip = 0; ip < n_in; ++ip) {
            const bound_state<S>& in = in_params[ip];
            const link* prev = step == 0 ? nullptr : &links[step - 1][param_to_link[step - 1][ip]];
            const uint32_t orig = prev ? prev->seed_idx : ip;
            const uint32_t skip = prev ? prev->n_skipped : 0u;
            const uint32_t consec = prev ? prev->n_consecutive_skipped : 0u;
            const double chi2_prev = prev ? prev->chi2_sum : 0.0;
            const uint32_t ndf_prev = prev ? prev->ndf_sum : 0u;

            const ckf_surface& sf = det.surfaces[in.surface];
            const uint32_t lo = in.surface == 0 ? 0u : meas_ranges[in.surface - 1];
            const uint32_t up = meas_ranges[in.surface];

            struct cand {
                link l;
                bound_state<S> p;
                S chi2;
            };
            std::vector<cand> best;
            if (mask) {
                // pruned run: the listed candidates in (chi2, index) order,
                // the kept ones updated until N succeed
                const bool rule0 = step == 0 && !(sf.thickness_x0 > 0.f);
                std::vector<std::pair<S, uint32_t>> listed;
                for (uint32_t mi = lo; mi < up; ++mi) {
                    const S chi2 = predicted_chi2<S>(load_meas<S>(meas, mi), in);
                    if (chi2 > chi2_max || chi2 < S(0)) continue;
                    if (rule0 && chi2 != S(0)) continue;
                    listed.push_back({chi2, mi});
                }
                std::sort(listed.begin(), listed.end());
                if (listed.size() > mask->list_len) listed.resize(mask->list_len);
                // the branch part of the candidates' verdict keys
                uint64_t branch = verdict_mix(verdict_mix(0u, orig), step);
                if (prev) {
                    link L = *prev;
                    for (;;) {
                        branch = verdict_mix(branch, L.meas_idx);
                        if (L.step == 0u) break;
                        L = back(L);
                    }
                }
                for (const auto& lc : listed) {
                    if (best.size() >= cfg.max_num_branches_per_surface) break;
                    if (!mask->has(branch, lc.second)) continue;
                    const meas_view<S> mv = load_meas<S>(meas, lc.second);
                    bound_state<S> filt;
                    int res = CKF_FIT_SUCCESS;
                    if (rule0) {
                        filt = in;
                        filt.cov.a[0][0] = mv.var[0];
                        filt.cov.a[1][1] = mv.var[1];
                    } else {
                        res = gain_update<S>(mv, in, filt);
                    }
                    if (res != CKF_FIT_SUCCESS) continue;
                    link l{step, ip, lc.second, orig, skip, 0u, static_cast<double>(lc.first),
                           chi2_prev + static_cast<double>(lc.first), ndf_prev + mv.dims};
                    if (!std::is_same<S, double>::value) {
                        l.chi2_sum = static_cast<double>(static_cast<S>(chi2_prev) + lc.first);
                    }
                    best.push_back({l, filt, lc.first});
                }
            }
            for (uint32_t mi = mask ? up : lo; mi < up; ++mi) {
                const meas_view<S> mv = load_meas<S>(meas, mi);
                const S chi2 = predicted_chi2<S>(mv, in);
                if (chi2 > chi2_max || chi2 < S(0)) continue;
                bound_state<S> filt;
                int res = CKF_FIT_ERROR_OTHER;
                if (step == 0 && !(sf.thickness_x0 > 0.f)) {
                    if (chi2 == S(0)) {
                        res = CKF_FIT_SUCCESS;
                        filt = in;
                        filt.cov.a[0][0] = mv.var[0];
                        filt.cov.a[1][1] = mv.var[1];
                    }
                } else {
                    res = gain_update<S>(mv, in, filt);
                }
                if (res == CKF_FIT_SUCCESS) {
                    link l{step, ip, mi, orig, skip, 0u, static_cast<double>(chi2),
                           chi2_prev + static_cast<double>(chi2), ndf_prev + mv.dims};
                    if (!std::is_same<S, double>::value) {
                        // FP32 build accumulates the chi2 sum in FP32, as traccc's scalar does
                        l.chi2_sum = static_cast<double>(static_cast<S>(chi2_prev) + chi2);
                    }
                    best.push_back({l, filt, chi2});
                }
            }
            std::sort(best.begin(), best.end(), [](const cand& a, const cand& b) {
                return a.chi2 < b.chi2 || (a.chi2 == b.chi2 && a.l.meas_idx < b.l.meas_idx);
            });
            const uint32_t nb = std::min<uint32_t>(cfg.max_num_branches_per_surface,
                                                   static_cast<uint32_t>(best.size()));
            for (uint32_t i = 0; i < nb; ++i) {
                links[step].push_back(best[i].l);
                updated.push_back(best[i].p);
            }
            if (nb == 0) {
                links[step].push_back({step, ip, std::numeric_limits<uint32_t>::max(), orig, skip + 1,
                                       consec + 1, static_cast<double>(std::numeric_limits<float>::max()),
                                       chi2_prev, ndf_prev});
                updated.push_back(in);
            }
        }

        // deduplication, reference :371-488
        const size_t n_links = links[step].size();
        stats.links += n_links;
        std::vector<uint8_t> alive(n_links, 1u);
        if (step >= cfg.duplicate_removal_minimum_length) {
            std::map<uint32_t, std::vector<size_t>> groups;
            for (size_t i = 0; i < n_links; ++i) {
                link L = links[step][i];
                while (L.meas_idx >= n_meas && L.step != 0u) L = back(L);
                groups[L.meas_idx].push_back(i);
            }
            for (const auto& g : groups) {
                const std::vector<size_t>& tr = g.second;
                for (size_t i = 0; i < tr.size(); ++i) {
                    const link& base = links[step][tr[i]];
                    // ndf_sum - 5 in unsigned arithmetic, as the reference writes it
                    const S p_this = prob<S>(static_cast<S>(base.chi2_sum),
                                             static_cast<S>(static_cast<uint32_t>(base.ndf_sum - 5u)));
                    if (step + 1 - base.n_skipped <= cfg.duplicate_removal_minimum_length ||
                        base.ndf_sum <= 5u)
                        continue;
                    for (size_t j = 0; j < tr.size(); ++j) {
                        if (i == j) continue;
                        link a = base;
                        link b = links[step][tr[j]];
                        if (step + 1 - b.n_skipped <= cfg.duplicate_removal_minimum_length ||
                            base.ndf_sum <= 5u)
                            continue;
                        bool dominated = true;
                        const S p_that = prob<S>(static_cast<S>(b.chi2_sum),
                                                 static_cast<S>(static_cast<uint32_t>(b.ndf_sum - 5u)));
                        while (true) {
                            while (a.meas_idx >= n_meas && a.step != 0u) a = back(a);
                            while (b.meas_idx >= n_meas && b.step != 0u) b = back(b);
                            if (a.meas_idx == b.meas_idx) {
                                if (a.step == 0u) break;
                                if (b.step == 0u) {
                                    dominated = false;
                                    break;
                                }
                                a = back(a);
                                b = back(b);
                            } else {
                                dominated = false;
                                break;
                            }
                        }
                        if (p_this != p_that) dominated = dominated && (p_that >= p_this);
                        else dominated = dominated && (tr[j] < tr[i]);
                        if (dominated) {
                            alive[tr[i]] = 0u;
                            break;
                        }
                    }
                }
            }
        }

        // propagation, reference :493-627
        for (uint32_t li = 0; li < n_links; ++li) {
            if (!alive[li]) continue;
            const link& L = links[step][li];
            if (++n_trks_per_seed[L.seed_idx] > cfg.max_num_branches_per_seed) continue;
            if (L.n_skipped > cfg.max_num_skipping_per_cand) {
                tips.push_back({step, li});
                continue;
            }
            if (L.n_consecutive_skipped > cfg.max_num_consecutive_skipped) {
                tips.push_back({step, li});
                continue;
            }
            bound_state<S> p = updated[li];
            // hypothesis charge follows the track's charge sign
            const S q_abs = abs_q;
            if (step == 0) add_material<S>(det.surfaces[p.surface], p, mass, q_abs);
            bound_state<S> o;
            ++stats.props;
            const int r = propagate_bound<S>(cfg, det, field, p, mass, q_abs, o, nullptr, nullptr);
            bool success = (r == 0);
            bool valid = success;
            if (success) {
                add_material<S>(det.surfaces[o.surface], o, mass, q_abs);
                const S th = o.vec.a[3][0];
                if (th <= S(0) || th >= static_cast<S>(2.0 * kPi)) valid = false;
                if (!std::isfinite(o.vec.a[2][0])) valid = false;
                if (std::fabs(o.vec.a[4][0]) == S(0)) valid = false;
                if (valid) {
                    out_params.push_back(o);
                    param_to_link[step].push_back(li);
                }
            }
            if (!valid && step >= cfg.min_track_candidates_per_track - 1u) tips.push_back({step, li});
            if (success && step == n_steps_max - 1u) tips.push_back({step, li});
        }
        in_params = std::move(out_params);
        out_params.clear();
    }
    stats.tips = tips.size();

    // build tracks, reference :637-715
    std::vector<track_out> result;
    result.reserve(tips.size());
    for (const auto& tip : tips) {
        link L = links[tip.first][tip.second];
        const S vote_pval = prob<S>(static_cast<S>(L.chi2_sum), static_cast<S>(L.ndf_sum) - S(5));
        const uint32_t n_cands = tip.first + 1 - L.n_skipped;
        if (n_cands < cfg.min_track_candidates_per_track || n_cands > cfg.max_track_candidates_per_track)
            continue;
        std::vector<uint32_t> ids(n_cands);
        S ndf = S(0), chi2 = S(0);
        bool complete = false;
        for (int k = static_cast<int>(n_cands) - 1; k >= 0; --k) {
            while (L.meas_idx >= n_meas && L.step != 0u) L = back(L);
            if (L.meas_idx >= n_meas) break;
            ids[k] = L.meas_idx;
            ndf += static_cast<S>(meas.dimensions[L.meas_idx]);
            chi2 += static_cast<S>(L.chi2);
            if (k == 0) {
                complete = true;
                break;
            }
            L = back(L);
        }
        if (!complete) continue;
        ndf -= S(5);
        track_out t;
        t.seed = L.seed_idx;
        t.ndf = static_cast<double>(ndf);
        t.chi2 = static_cast<double>(chi2);
        t.pval = static_cast<double>(prob<S>(chi2, ndf));
        t.nholes = L.n_skipped;
        t.vote_pval = static_cast<double>(vote_pval);
        t.meas = std::move(ids);
        result.push_back(std::move(t));
    }

    // Measurement-vote pruning (the reference runs it on the device chain only):
    // device/common/include/traccc/finding/device/impl/gather_best_tips_per_measurement.ipp:27-180,
    // gather_measurement_votes.ipp:19-47, update_tip_length_buffer.ipp:20-54, driven by
    // device/common/src/finding/combinatorial_kalman_filter_algorithm.cpp:568-634.
    // Every measurement keeps the max_num_tracks_per_measurement tracks with the
    // highest p-value of their tip's running chi2 / ndf sums; a track survives
    // when the measurements that kept it are at least
    // min_measurement_voting_fraction of its measurements.  Equal p-values rank
    // by lower track index (the reference leaves that order to its insertion
    // mutex); the survivors keep their order.
    if (cfg.max_num_tracks_per_measurement > 0u && !result.empty()) {
        const uint32_t K = cfg.max_num_tracks_per_measurement;
        std::vector<std::vector<std::pair<double, uint32_t>>> per_meas(n_meas);
        for (uint32_t i = 0; i < result.size(); ++i) {
            for (uint32_t m : result[i].meas) per_meas[m].push_back({result[i].vote_pval, i});
        }
        std::vector<uint32_t> votes(result.size(), 0u);
        for (auto& v : per_meas) {
            std::sort(v.begin(), v.end(), [](const std::pair<double, uint32_t>& a,
                                             const std::pair<double, uint32_t>& b) {
                return a.first > b.first || (a.first == b.first && a.second < b.second);
            });
            for (size_t k = 0; k < v.size() && k < K; ++k) ++votes[v[k].second];
        }
        std::vector<track_out> kept;
        kept.reserve(result.size());
        for (uint32_t i = 0; i < result.size(); ++i) {
            const float frac = static_cast<float>(votes[i]) /
                               static_cast<float>(result[i].meas.size());
            if (!(frac < cfg.min_measurement_voting_fraction)) kept.push_back(std::move(result[i]));
        }
        result = std::move(kept);
    }
    return result;
}

// ---------------------------------------------------------------------------
// Progressive Kalman filter (finding_config.run_pkf): one branch per seed that
// takes the best measurement of every sensitive surface it reaches.
//   run_progressive_kalman_filter  core/include/traccc/finding/details/run_progressive_kalman_filter.hpp:55-258
//   progressive_kalman_filter      core/include/traccc/finding/details/progressive_kalman_filter.hpp:53-209
//   measurement_updater actor      core/include/traccc/finding/actors/measurement_kalman_updater.hpp:94-260
//   find_optimal_measurement       core/include/traccc/finding/measurement_selector.hpp:224-264
// On every surface (the seed's included) the material noise comes first, then
// the update on the lowest-chi2 measurement (first index on ties, the first
// one at or below FLT_EPSILON ends the search); a chi2 above chi2_max or a
// failed update is a hole.  The walk ends (a finished track) on too many
// holes / consecutive holes, on max_track_candidates_per_track states or when
// the navigation leaves the detector; it fails on an aborted propagation or a
// total path beyond path_limit.  Every seed yields one entry; a failed or too
// short track is an empty entry (as the reference pushes a default track).
// ---------------------------------------------------------------------------
constexpr uint32_t kNoSeed = 0xffffffffu;

template <typename S>
std::vector<track_out> pkf(const ckf_finding_config& cfg, const ckf_detector& det,
                           const ckf_bfield& field, const ckf_measurements& meas,
                           const ckf_bound_params& seeds, run_stats& stats) {
    const uint32_t n_meas = meas.size;
    const S mass = static_cast<S>(cfg.ptc_mass);
    const S abs_q = static_cast<S>(std::fabs(cfg.ptc_charge));
    const S chi2_max = static_cast<S>(cfg.chi2_max);
    const S eps = static_cast<S>(std::numeric_limits<float>::epsilon());
    std::vector<uint32_t> meas_ranges(det.n_surfaces);
    for (uint32_t s = 0; s < det.n_surfaces; ++s) {
        if (!det.surfaces[s].sensitive) {
            meas_ranges[s] = s == 0 ? 0u : meas_ranges[s - 1];
            continue;
        }
        meas_ranges[s] = static_cast<uint32_t>(
            std::upper_bound(meas.surface_link, meas.surface_link + n_meas,
                             det.surfaces[s].identifier) - meas.surface_link);
    }
    // legs between surfaces: no minimum step and no per-leg step budget (the
    // PKF propagation has no ckf_aborter); the path limit holds for the walk
    ckf_finding_config leg = cfg;
    leg.min_step_length_for_next_surface = 0.f;
    leg.max_step_counts_for_next_surface = 0x7fffffffu;
    auto momentum_ok = [&](const bound_state<S>& b, S div) {
        const S p = abs_q / std::fabs(b.vec.a[4][0]);
        const S pT = p * std::sin(b.vec.a[3][0]);
        return pT > static_cast<S>(cfg.min_pT) / div && p > static_cast<S>(cfg.min_p) / div;
    };

    std::vector<track_out> result(seeds.size);
    for (uint32_t si = 0; si < seeds.size; ++si) {
        track_out& t = result[si];
        t.seed = kNoSeed;
        bound_state<S> b;
        b.surface = surface_index(det, seeds.surface_link[si]);
        for (int k = 0; k < 6; ++k) b.vec.a[k][0] = static_cast<S>(seeds.vector[6 * si + k]);
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) b.cov.a[r][c] = static_cast<S>(seeds.covariance[36 * si + 6 * r + c]);
        if (!momentum_ok(b, S(10))) continue;
        std::vector<uint32_t> states;
        uint32_t holes = 0u, consec = 0u, ndf_sum = 0u;
        S chi2_sum = S(0), path = S(0);
        bool good = false;
        for (;;) {
            const ckf_surface& sf = det.surfaces[b.surface];
            add_material<S>(sf, b, mass, abs_q);
            // find_optimal_measurement
            const uint32_t lo = b.surface == 0 ? 0u : meas_ranges[b.surface - 1];
            const uint32_t up = meas_ranges[b.surface];
            uint32_t best = kNoSeed;
            S best_chi2 = std::numeric_limits<S>::max();
            for (uint32_t mi = lo; mi < up; ++mi) {
                const S chi2 = predicted_chi2<S>(load_meas<S>(meas, mi), b);
                if (chi2 < best_chi2 && chi2 >= S(0)) {
                    best = mi;
                    best_chi2 = chi2;
                    if (best_chi2 <= eps) break;
                }
            }
            int res = CKF_FIT_ERROR_OTHER;
            if (best != kNoSeed && best_chi2 <= chi2_max) {
                const meas_view<S> mv = load_meas<S>(meas, best);
                if (best_chi2 <= eps && !(sf.thickness_x0 > 0.f)) {
                    b.cov.a[0][0] = mv.var[0];
                    b.cov.a[1][1] = mv.var[1];
                    res = CKF_FIT_SUCCESS;
                } else {
                    bound_state<S> f;
                    res = gain_update<S>(mv, b, f);
                    if (res == CKF_FIT_SUCCESS) b = f;
                }
            }
            bool exit = false;
            if (res != CKF_FIT_SUCCESS) {
                ++holes;
                ++consec;
                exit = holes > cfg.max_num_skipping_per_cand || consec > cfg.max_num_consecutive_skipped;
            } else {
                states.push_back(best);
                consec = 0u;
                ndf_sum += meas.dimensions[best];
                chi2_sum += best_chi2;
                exit = states.size() >= cfg.max_track_candidates_per_track;
            }
            if (exit) {
                good = true;
                break;
            }
            bound_state<S> o;
            S leg_path = S(0);
            ++stats.props;
            const int r = propagate_bound<S>(leg, det, field, b, mass, abs_q, o, &leg_path, nullptr);
            path += leg_path;
            if (r == 1) {
                good = true;  // left the detector
                break;
            }
            if (r != 0 || std::fabs(path) > static_cast<S>(cfg.propagation.path_limit)) break;
            b = o;
        }
        stats.steps = std::max<uint32_t>(stats.steps, static_cast<uint32_t>(states.size() + holes));
        if (!good || !momentum_ok(b, S(1)) || states.size() < cfg.min_track_candidates_per_track ||
            static_cast<int>(ndf_sum) - 5 < 0) {
            continue;
        }
        const S ndf = static_cast<S>(static_cast<int>(ndf_sum) - 5);
        t.seed = si;
        t.ndf = static_cast<double>(ndf);
        t.chi2 = static_cast<double>(chi2_sum);
        t.pval = static_cast<double>(prob<S>(chi2_sum, ndf));
        t.vote_pval = t.pval;
        t.nholes = holes;
        t.meas = std::move(states);
    }
    stats.tips = seeds.size;
    return result;
}

// ---------------------------------------------------------------------------
// Kalman fit with the two-filter smoother (kalman_fitting_algorithm)
//   kalman_fitter::fit / fit_iteration / filter / smooth
//                      core/include/traccc/fitting/kalman_filter/kalman_fitter.hpp:177-420
//   update_statistics / check_fitting_result                        :419-550
//   kalman_actor       core/include/traccc/fitting/kalman_filter/kalman_actor.hpp (forward
//                      gain update, hole counting, surface sequence, backward smoothing)
//   two_filters_smoother core/include/traccc/fitting/kalman_filter/two_filters_smoother.hpp:38-294
//   inflate_covariance core/include/traccc/edm/track_parameters.hpp:93-106
// Forward: the covariance-inflated seed goes through every sensitive surface
// the navigation reaches; the material of a surface comes first, then the
// gain update with the candidate's measurement of that surface; any other
// sensitive surface is a hole (kept in the sequence when it has material).
// Backward: from the last state (filtered, inflated) along the recorded
// sequence with direct transport; on each state the smoother (information
// form, as the reference), then the backward filter, then the material.
// Deliberate deviation, applied identically in the product: before the two
// estimates are combined, the backward prediction's phi is moved by 2 pi to
// within pi of the forward one (the reference combines the raw angles, which
// is wrong for tracks at phi ~ +-pi).
// A navigation that leaves the detector before the last state leaves states
// unfiltered; the reference's smoother then fails on the unfiltered state
// (FAILURE_SMOOTHER), which is the outcome given here directly.
// ---------------------------------------------------------------------------
template <typename S, int N>
bool invert(const mat<S, N, N>& in, mat<S, N, N>& out) {
    mat<S, N, N> a = in;
    out = mat<S, N, N>::identity();
    for (int c = 0; c < N; ++c) {
        int piv = c;
        for (int r = c + 1; r < N; ++r)
            if (std::fabs(a.a[r][c]) > std::fabs(a.a[piv][c])) piv = r;
        if (!(std::fabs(a.a[piv][c]) > S(0))) return false;
        if (piv != c) {
            for (int k = 0; k < N; ++k) {
                std::swap(a.a[c][k], a.a[piv][k]);
                std::swap(out.a[c][k], out.a[piv][k]);
            }
        }
        const S d = a.a[c][c];
        for (int k = 0; k < N; ++k) {
            a.a[c][k] /= d;
            out.a[c][k] /= d;
        }
        for (int r = 0; r < N; ++r) {
            if (r == c) continue;
            const S f = a.a[r][c];
            if (f == S(0)) continue;
            for (int k = 0; k < N; ++k) {
                a.a[r][k] -= f * a.a[c][k];
                out.a[r][k] -= f * out.a[c][k];
            }
        }
    }
    return true;
}

template <typename S>
void inflate(mat<S, 6, 6>& c, S f) {
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) c.a[i][j] = (i == j) ? c.a[i][i] * f : S(0);
}

template <typename S>
bool regularize(mat<S, 6, 6>& c) {
    for (int i = 0; i < 6; ++i)
        if (c.a[i][i] < S(-0.01)) return false;
    for (int i = 0; i < 6; ++i) c.a[i][i] = std::fabs(c.a[i][i]);
    return true;
}

// kalman_fitter_status values missing from ckf_fitter_status
enum : int {
    kFitSmootherChi2Negative = 4,
    kFitSmootherChi2NotFinite = 5,
    kFitSmootherInvalidCov = 6,
    kFitSmootherSkipped = 7,
    kFitUpdaterSkipped = 9,
    kFitSequenceOverflow = 10,
    kFitPropagationFailure = 11
};

template <typename S>
struct fit_state {
    uint8_t flags = 0u;
    S filtered_chi2 = S(0), smoothed_chi2 = S(0), backward_chi2 = S(0);
    bool filtered_ok = false;
    bound_state<S> filtered{}, smoothed{};
    uint32_t surface = 0u;
};

template <typename S>
struct fit_result {
    uint16_t outcome = CKF_OUTCOME_UNKNOWN;
    bound_state<S> params{};
    S ndf = S(0), chi2 = S(0), pval = S(0);
    uint32_t nholes = 0u;
    std::vector<fit_state<S>> states;
};

// two_filters_smoother::operator() on state st with the backward prediction
// b; b becomes the backward filtered state.
template <typename S>
int smooth_state(fit_state<S>& st, const meas_view<S>& m, bound_state<S>& b) {
    const mat<S, 6, 6>& Cf = st.filtered.cov;
    mat<S, 6, 1> xp = b.vec;
    const S twopi = static_cast<S>(2.0 * kPi);
    const S dphi = xp.a[2][0] - st.filtered.vec.a[2][0];
    if (dphi > static_cast<S>(kPi)) xp.a[2][0] -= twopi;
    if (dphi < -static_cast<S>(kPi)) xp.a[2][0] += twopi;
    mat<S, 6, 6> Cp_inv, Cf_inv, Cs;
    if (!invert<S, 6>(b.cov, Cp_inv) || !invert<S, 6>(Cf, Cf_inv)) return CKF_FIT_ERROR_INVERSION;
    if (!invert<S, 6>(Cp_inv + Cf_inv, Cs)) return CKF_FIT_ERROR_INVERSION;
    if (!regularize(Cs)) return kFitSmootherInvalidCov;
    mat<S, 6, 1> xs = Cs * (Cf_inv * st.filtered.vec + Cp_inv * xp);
    if (!std::isfinite(xs.a[3][0]) || !std::isfinite(xs.a[2][0])) return CKF_FIT_ERROR_INVERSION;
    if (std::fabs(xs.a[4][0]) == S(0)) return CKF_FIT_ERROR_QOP_ZERO;
    if (!wrap_angles(xs)) return CKF_FIT_ERROR_THETA_POLE;
    const mat<S, 2, 6> H = projector(m);
    mat<S, 2, 2> V = mat<S, 2, 2>::zero();
    V.a[0][0] = m.var[0];
    V.a[1][1] = m.var[1];
    mat<S, 2, 1> ml;
    ml.a[0][0] = m.pos[0];
    ml.a[1][0] = m.dims == 2u ? m.pos[1] : S(0);
    const mat<S, 2, 1> rs = ml - H * xs;
    const mat<S, 2, 2> Rs = V - H * (Cs * transpose(H));
    const S chi2_s = (transpose(rs) * masked_inv(Rs, m.dims) * rs).a[0][0];
    if (chi2_s < S(-10)) return kFitSmootherChi2Negative;
    if (!std::isfinite(chi2_s)) return kFitSmootherChi2NotFinite;
    // backward filter (the reference's inline gain update)
    const mat<S, 6, 2> PHt = b.cov * transpose(H);
    const mat<S, 6, 2> K = PHt * masked_inv(H * PHt + V, m.dims);
    mat<S, 6, 1> xb = b.vec + K * (ml - H * b.vec);
    if (!std::isfinite(xb.a[3][0]) || !std::isfinite(xb.a[2][0])) return CKF_FIT_ERROR_INVERSION;
    if (std::fabs(xb.a[4][0]) == S(0)) return CKF_FIT_ERROR_QOP_ZERO;
    if (!wrap_angles(xb)) return CKF_FIT_ERROR_THETA_POLE;
    const mat<S, 6, 6> IKH = mat<S, 6, 6>::identity() - K * H;
    mat<S, 6, 6> Cb = IKH * b.cov * transpose(IKH) + K * V * transpose(K);
    if (!regularize(Cb)) return kFitSmootherInvalidCov;
    const mat<S, 2, 1> r = ml - H * xb;
    const mat<S, 2, 2> R = (mat<S, 2, 2>::identity() - H * K) * V;
    const S chi2_b = (transpose(r) * masked_inv(R, m.dims) * r).a[0][0];
    if (chi2_b < S(0)) return kFitSmootherChi2Negative;
    if (!std::isfinite(chi2_b)) return kFitSmootherChi2NotFinite;
    st.smoothed.surface = st.surface;
    st.smoothed.vec = xs;
    st.smoothed.cov = Cs;
    st.smoothed_chi2 = chi2_s;
    st.backward_chi2 = chi2_b;
    st.flags |= CKF_STATE_SMOOTHED;
    b.vec = xb;
    b.cov = Cb;
    return CKF_FIT_SUCCESS;
}

template <typename S>
fit_result<S> kalman_fit(const ckf_fitting_config& fc, const ckf_detector& det, const ckf_bfield& field,
                         const ckf_measurements& meas, const bound_state<S>& seed,
                         const uint32_t* idx, uint32_t n, const std::vector<uint32_t>& meas_sf) {
    // navigation settings of the fit (no minimum step: no ckf_aborter)
    ckf_finding_config nc{};
    nc.propagation = fc.propagation;
    nc.min_step_length_for_next_surface = 0.f;
    nc.max_step_counts_for_next_surface = fc.max_step_counts;
    nc.min_p = fc.min_p;
    nc.min_pT = fc.min_pT;
    nc.ptc_mass = fc.ptc_mass;
    nc.ptc_charge = fc.ptc_charge;
    const S mass = static_cast<S>(fc.ptc_mass);
    const S abs_q = static_cast<S>(std::fabs(fc.ptc_charge));
    const S infl = static_cast<S>(fc.covariance_inflation_factor);
    const size_t cap = std::max<size_t>(size_t(n) * fc.surface_sequence_size_factor,
                                        fc.min_surface_sequence_capacity);
    fit_result<S> R;
    R.states.resize(n);
    for (uint32_t k = 0; k < n; ++k) R.states[k].surface = meas_sf[idx[k]];
    bound_state<S> start = seed;
    for (uint32_t it = 0; it < std::max(fc.n_iterations, 1u); ++it) {
        for (auto& st : R.states) {
            st.flags = 0u;
            st.filtered_ok = false;
        }
        // ---- forward filter
        bound_state<S> b = start;
        inflate(b.cov, infl);
        uint32_t k = 0u, holes = 0u;
        std::vector<int64_t> seq;  // >= 0: state index; < 0: -(surface + 1) of a hole
        int fw = CKF_FIT_SUCCESS;
        bool complete = false;
        for (;;) {
            add_material<S>(det.surfaces[b.surface], b, mass, abs_q);
            if (b.surface == R.states[k].surface) {
                const meas_view<S> mv = load_meas<S>(meas, idx[k]);
                bound_state<S> f;
                const int r = gain_update<S>(mv, b, f);
                if (r != CKF_FIT_SUCCESS) {
                    fw = r;
                    break;
                }
                R.states[k].filtered = f;
                R.states[k].filtered_ok = true;
                R.states[k].filtered_chi2 = predicted_chi2<S>(mv, f);
                b = f;
                if (seq.size() == cap) {
                    fw = kFitSequenceOverflow;
                    break;
                }
                seq.push_back(k);
                if (++k == n) {
                    complete = true;
                    break;
                }
            } else {
                bool later = false;
                for (uint32_t j = k + 1; j < n; ++j) later = later || R.states[j].surface == b.surface;
                if (later) {
                    fw = kFitUpdaterSkipped;
                    break;
                }
                ++holes;
                if (det.surfaces[b.surface].thickness_x0 > 0.f) {
                    if (seq.size() == cap) {
                        fw = kFitSequenceOverflow;
                        break;
                    }
                    seq.push_back(-static_cast<int64_t>(b.surface) - 1);
                }
            }
            bound_state<S> o;
            // the fit reaches a module when its crossing lies within
            // min_mask_tolerance of the mask
            const int r = propagate_bound<S>(nc, det, field, b, mass, abs_q, o, nullptr, nullptr,
                                             static_cast<S>(nc.propagation.min_mask_tolerance));
            if (r == 1) break;  // left the detector: states left unfiltered
            if (r != 0) {
                fw = kFitPropagationFailure;
                break;
            }
            b = o;
        }
        R.nholes = holes;
        if (fw != CKF_FIT_SUCCESS) {
            R.outcome = fw == kFitUpdaterSkipped       ? CKF_OUTCOME_FAILURE_NOT_ALL_FITTED
                        : fw == kFitPropagationFailure ? CKF_OUTCOME_FAILURE_FORWARD_PROPAGATION
                        : fw == kFitSequenceOverflow   ? CKF_OUTCOME_FAILURE_SMOOTHER
                                                       : CKF_OUTCOME_FAILURE_FITTER;
            return R;
        }
        if (!complete) {
            R.outcome = CKF_OUTCOME_FAILURE_SMOOTHER;
            return R;
        }
        // ---- backward smoother
        fit_state<S>& last = R.states[n - 1];
        last.smoothed = last.filtered;
        bound_state<S> bb = last.filtered;
        inflate(bb.cov, infl);
        int bw = CKF_FIT_SUCCESS;
        for (size_t e = seq.size(); e-- > 0;) {
            const uint32_t target = seq[e] >= 0 ? R.states[seq[e]].surface
                                                : static_cast<uint32_t>(-seq[e] - 1);
            if (target != bb.surface) {
                bound_state<S> o;
                if (propagate_direct<S>(nc, det, field, bb, target, mass, abs_q, o) != 0) {
                    bw = kFitPropagationFailure;
                    break;
                }
                bb = o;
            }
            if (seq[e] >= 0) {
                const uint32_t j = static_cast<uint32_t>(seq[e]);
                const int r = smooth_state<S>(R.states[j], load_meas<S>(meas, idx[j]), bb);
                if (r != CKF_FIT_SUCCESS) {
                    bw = r;
                    break;
                }
            }
            add_material<S>(det.surfaces[target], bb, mass, abs_q);
        }
        if (bw != CKF_FIT_SUCCESS) {
            R.outcome = bw == kFitSmootherSkipped      ? CKF_OUTCOME_FAILURE_NOT_ALL_SMOOTHED
                        : bw == kFitPropagationFailure ? CKF_OUTCOME_FAILURE_BACKWARD_PROPAGATION
                                                       : CKF_OUTCOME_FAILURE_SMOOTHER;
            return R;
        }
        // ---- statistics and outcome
        S ndf = S(0), chi2 = S(0);
        bool have_params = false;
        for (uint32_t j = 0; j < n; ++j) {
            const fit_state<S>& st = R.states[j];
            if (st.flags & CKF_STATE_SMOOTHED) {
                if (!have_params) {
                    R.params = st.smoothed;
                    have_params = true;
                }
                ndf += static_cast<S>(meas.dimensions[idx[j]]);
                chi2 += st.backward_chi2;
            }
        }
        ndf -= S(5);
        R.ndf = ndf;
        R.chi2 = chi2;
        R.pval = prob<S>(chi2, ndf);
        R.outcome = ndf > S(0) ? CKF_OUTCOME_SUCCESS : CKF_OUTCOME_FAILURE_NON_POSITIVE_NDF;
        // the next iteration starts from the first state's filtered parameters
        start = R.states[0].filtered;
    }
    return R;
}

template <typename T>
T* alloc_n(size_t n) {
    return static_cast<T*>(std::calloc(n ? n : 1, sizeof(T)));
}

void fill_out(const std::vector<track_out>& tr, const ckf_bound_params& seeds, const run_stats& st,
              oracle_tracks* out) {
    std::memset(out, 0, sizeof(*out));
    const uint32_t n = static_cast<uint32_t>(tr.size());
    uint32_t nc = 0;
    for (const auto& t : tr) nc += static_cast<uint32_t>(t.meas.size());
    out->n_tracks = n;
    out->n_constituents = nc;
    out->params_surface_link = alloc_n<uint64_t>(n);
    out->params_vector = alloc_n<float>(6u * n);
    out->params_covariance = alloc_n<float>(36u * n);
    out->ndf = alloc_n<double>(n);
    out->chi2 = alloc_n<double>(n);
    out->pval = alloc_n<double>(n);
    out->nholes = alloc_n<uint32_t>(n);
    out->constituent_offsets = alloc_n<uint32_t>(n + 1u);
    out->constituent_index = alloc_n<uint32_t>(nc);
    uint32_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const track_out& t = tr[i];
        if (t.seed == kNoSeed) {
            // an empty entry (progressive filter without a track): default parameters
            out->params_surface_link[i] = ~0ull;
            out->constituent_offsets[i] = off;
            continue;
        }
        out->params_surface_link[i] = seeds.surface_link[t.seed];
        std::memcpy(out->params_vector + 6u * i, seeds.vector + 6u * t.seed, 6u * sizeof(float));
        std::memcpy(out->params_covariance + 36u * i, seeds.covariance + 36u * t.seed, 36u * sizeof(float));
        out->ndf[i] = t.ndf;
        out->chi2[i] = t.chi2;
        out->pval[i] = t.pval;
        out->nholes[i] = t.nholes;
        out->constituent_offsets[i] = off;
        for (uint32_t m : t.meas) out->constituent_index[off++] = m;
    }
    out->constituent_offsets[n] = off;
    out->n_steps = st.steps;
    out->n_links = st.links;
    out->n_propagations = st.props;
    out->n_tips = st.tips;
}

}  // namespace oracle

extern "C" {

int oracle_ckf_run(const ckf_finding_config* cfg, const ckf_detector* det, const ckf_bfield* field,
                   const ckf_measurements* meas, const ckf_bound_params* seeds, int precision,
                   oracle_tracks* out) {
    if (!cfg || !det || !field || !meas || !seeds || !out) return CKF_ERR_INVALID_ARGUMENT;
    if (cfg->min_track_candidates_per_track == 0u) return CKF_ERR_INVALID_ARGUMENT;
    oracle::run_stats st;
    // the operator's dispatch: core/src/finding/combinatorial_kalman_filter_algorithm.cpp:47-57
    const bool use_pkf = cfg->run_pkf && (cfg->max_num_branches_per_seed == 1u ||
                                          cfg->max_num_branches_per_surface == 1u);
    std::vector<oracle::track_out> tr;
    if (use_pkf) {
        tr = precision == 32 ? oracle::pkf<float>(*cfg, *det, *field, *meas, *seeds, st)
                             : oracle::pkf<double>(*cfg, *det, *field, *meas, *seeds, st);
    } else {
        tr = precision == 32 ? oracle::ckf<float>(*cfg, *det, *field, *meas, *seeds, st)
                             : oracle::ckf<double>(*cfg, *det, *field, *meas, *seeds, st);
    }
    oracle::fill_out(tr, *seeds, st, out);
    return CKF_OK;
}

void oracle_tracks_free(oracle_tracks* t) {
    if (!t) return;
    std::free(t->params_surface_link);
    std::free(t->params_vector);
    std::free(t->params_covariance);
    std::free(t->ndf);
    std::free(t->chi2);
    std::free(t->pval);
    std::free(t->nholes);
    std::free(t->constituent_offsets);
    std::free(t->constituent_index);
    std::memset(t, 0, sizeof(*t));
}

int oracle_ckf_run_masked(const ckf_finding_config* cfg, const ckf_detector* det,
                          const ckf_bfield* field, const ckf_measurements* meas,
                          const ckf_bound_params* seeds, int precision, uint32_t list_len,
                          const uint64_t* kept, uint64_t n_kept, oracle_tracks* out) {
    const bool keep_all = kept == nullptr && n_kept == ~uint64_t(0);
    if (!cfg || !det || !field || !meas || !seeds || !out || (n_kept && !kept && !keep_all))
        return CKF_ERR_INVALID_ARGUMENT;
    if (cfg->min_track_candidates_per_track == 0u || cfg->run_pkf) return CKF_ERR_INVALID_ARGUMENT;
    for (uint64_t k = 1; !keep_all && k < n_kept; ++k) {
        if (kept[k - 1] > kept[k]) return CKF_ERR_INVALID_ARGUMENT;  // must be sorted
    }
    oracle::cand_mask mk;
    mk.kept = kept;
    mk.n_kept = keep_all ? 0u : static_cast<size_t>(n_kept);
    mk.keep_all = keep_all;
    mk.list_len = list_len;
    oracle::run_stats st;
    std::vector<oracle::track_out> tr =
        precision == 32 ? oracle::ckf<float>(*cfg, *det, *field, *meas, *seeds, st, &mk)
                        : oracle::ckf<double>(*cfg, *det, *field, *meas, *seeds, st, &mk);
    oracle::fill_out(tr, *seeds, st, out);
    return CKF_OK;
}

int oracle_ckf_run_many(const ckf_finding_config* cfg, const ckf_detector* det,
                        const ckf_bfield* field, const ckf_measurements* meas,
                        const ckf_bound_params* seeds, uint32_t n_events, int precision,
                        uint32_t n_threads, uint64_t* n_tracks_total, double* seconds) {
    if (!cfg || !det || !field || !meas || !seeds || n_threads == 0u) return CKF_ERR_INVALID_ARGUMENT;
    std::atomic<uint32_t> next{0u};
    std::atomic<uint64_t> total{0u};
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < n_threads; ++t) {
        pool.emplace_back([&]() {
            for (;;) {
                const uint32_t e = next.fetch_add(1u);
                if (e >= n_events) break;
                oracle::run_stats st;
                const size_t n = precision == 32
                                     ? oracle::ckf<float>(*cfg, *det, *field, meas[e], seeds[e], st).size()
                                     : oracle::ckf<double>(*cfg, *det, *field, meas[e], seeds[e], st).size();
                total.fetch_add(n);
            }
        });
    }
    for (auto& th : pool) th.join();
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (n_tracks_total) *n_tracks_total = total.load();
    return CKF_OK;
}

void oracle_fitted_free(oracle_fitted* f) {
    if (!f) return;
    void* ps[] = {f->outcome, f->params_surface, f->params_vector, f->params_covariance, f->ndf,
                  f->chi2, f->pval, f->nholes, f->state_offsets, f->state_flags, f->filtered_chi2,
                  f->smoothed_chi2, f->backward_chi2, f->filtered_vector, f->filtered_covariance,
                  f->smoothed_vector, f->smoothed_covariance};
    for (void* p : ps) std::free(p);
    std::memset(f, 0, sizeof(*f));
}

int oracle_fit_run(const ckf_fitting_config* cfg, const ckf_detector* det, const ckf_bfield* field,
                   const ckf_measurements* meas, uint32_t n_tracks, const uint64_t* seed_surface_link,
                   const float* seed_vec, const float* seed_cov, const uint32_t* offsets,
                   const uint32_t* meas_index, int precision, uint32_t n_threads, oracle_fitted* out,
                   double* seconds) {
    using namespace oracle;
    if (!cfg || !det || !field || !meas || !out || n_threads == 0u ||
        (n_tracks > 0u && (!seed_surface_link || !seed_vec || !seed_cov || !offsets || !meas_index))) {
        return CKF_ERR_INVALID_ARGUMENT;
    }
    std::memset(out, 0, sizeof(*out));
    // surface index of every measurement
    std::vector<uint32_t> meas_sf(meas->size, 0u);
    for (uint32_t i = 0; i < meas->size; ++i) meas_sf[i] = surface_index(*det, meas->surface_link[i]);
    const uint32_t ns = n_tracks ? offsets[n_tracks] : 0u;
    out->n_tracks = n_tracks;
    out->n_states = ns;
    out->outcome = alloc_n<uint16_t>(n_tracks);
    out->params_surface = alloc_n<uint32_t>(n_tracks);
    out->params_vector = alloc_n<double>(6u * n_tracks);
    out->params_covariance = alloc_n<double>(36u * n_tracks);
    out->ndf = alloc_n<double>(n_tracks);
    out->chi2 = alloc_n<double>(n_tracks);
    out->pval = alloc_n<double>(n_tracks);
    out->nholes = alloc_n<uint32_t>(n_tracks);
    out->state_offsets = alloc_n<uint32_t>(n_tracks + 1u);
    out->state_flags = alloc_n<uint8_t>(ns);
    out->filtered_chi2 = alloc_n<double>(ns);
    out->smoothed_chi2 = alloc_n<double>(ns);
    out->backward_chi2 = alloc_n<double>(ns);
    out->filtered_vector = alloc_n<double>(6u * size_t(ns));
    out->filtered_covariance = alloc_n<double>(36u * size_t(ns));
    out->smoothed_vector = alloc_n<double>(6u * size_t(ns));
    out->smoothed_covariance = alloc_n<double>(36u * size_t(ns));
    for (uint32_t t = 0; t <= n_tracks && n_tracks; ++t) out->state_offsets[t] = offsets[t];
    auto fit_one = [&](uint32_t t, auto tag) {
        using S = decltype(tag);
        bound_state<S> b;
        b.surface = surface_index(*det, seed_surface_link[t]);
        for (int k = 0; k < 6; ++k) b.vec.a[k][0] = static_cast<S>(seed_vec[6u * t + k]);
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) b.cov.a[r][c] = static_cast<S>(seed_cov[36u * t + 6 * r + c]);
        const uint32_t s0 = offsets[t], n = offsets[t + 1] - offsets[t];
        const fit_result<S> R = kalman_fit<S>(*cfg, *det, *field, *meas, b, meas_index + s0, n, meas_sf);
        out->outcome[t] = R.outcome;
        out->params_surface[t] = R.params.surface;
        for (int k = 0; k < 6; ++k) out->params_vector[6u * t + k] = static_cast<double>(R.params.vec.a[k][0]);
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c)
                out->params_covariance[36u * t + 6 * r + c] = static_cast<double>(R.params.cov.a[r][c]);
        out->ndf[t] = static_cast<double>(R.ndf);
        out->chi2[t] = static_cast<double>(R.chi2);
        out->pval[t] = static_cast<double>(R.pval);
        out->nholes[t] = R.nholes;
        for (uint32_t k = 0; k < n; ++k) {
            const fit_state<S>& st = R.states[k];
            const size_t g = s0 + k;
            out->state_flags[g] = st.flags;
            out->filtered_chi2[g] = static_cast<double>(st.filtered_chi2);
            out->smoothed_chi2[g] = static_cast<double>(st.smoothed_chi2);
            out->backward_chi2[g] = static_cast<double>(st.backward_chi2);
            for (int q = 0; q < 6; ++q) {
                out->filtered_vector[6u * g + q] = static_cast<double>(st.filtered.vec.a[q][0]);
                out->smoothed_vector[6u * g + q] = static_cast<double>(st.smoothed.vec.a[q][0]);
            }
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    out->filtered_covariance[36u * g + 6 * r + c] = static_cast<double>(st.filtered.cov.a[r][c]);
                    out->smoothed_covariance[36u * g + 6 * r + c] = static_cast<double>(st.smoothed.cov.a[r][c]);
                }
        }
    };
    std::atomic<uint32_t> next{0u};
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (uint32_t th = 0; th < n_threads; ++th) {
        pool.emplace_back([&]() {
            for (;;) {
                const uint32_t t0i = next.fetch_add(64u);
                if (t0i >= n_tracks) break;
                const uint32_t t1 = std::min(n_tracks, t0i + 64u);
                for (uint32_t t = t0i; t < t1; ++t) {
                    if (precision == 32) fit_one(t, 0.f);
                    else fit_one(t, 0.0);
                }
            }
        });
    }
    for (auto& th : pool) th.join();
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return CKF_OK;
}

double oracle_prob(double chi2, double ndf) { return oracle::prob<double>(chi2, ndf); }
float oracle_prob_f(float chi2, float ndf) { return oracle::prob<float>(chi2, ndf); }

int oracle_kf_update(const double vec[6], const double cov[36], const float m[2], const float var[2],
                     uint32_t dims, const uint8_t subspace[2], double* chi2, double out_vec[6],
                     double out_cov[36]) {
    using namespace oracle;
    bound_state<double> in, f;
    in.surface = 0;
    for (int i = 0; i < 6; ++i) in.vec.a[i][0] = vec[i];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) in.cov.a[r][c] = cov[6 * r + c];
    ckf_measurements M{};
    uint32_t d = dims;
    M.size = 1;
    M.local_position = m;
    M.local_variance = var;
    M.dimensions = &d;
    M.subspace = subspace;
    const meas_view<double> mv = load_meas<double>(M, 0);
    *chi2 = predicted_chi2<double>(mv, in);
    const int r = gain_update<double>(mv, in, f);
    if (r == CKF_FIT_SUCCESS) {
        for (int i = 0; i < 6; ++i) out_vec[i] = f.vec.a[i][0];
        for (int rr = 0; rr < 6; ++rr)
            for (int c = 0; c < 6; ++c) out_cov[6 * rr + c] = f.cov.a[rr][c];
    }
    return r;
}

int oracle_propagate(const ckf_finding_config* cfg, const ckf_detector* det, const ckf_bfield* field,
                     uint32_t start, const double vec[6], const double cov[36], int apply_start_material,
                     uint32_t* end_surface, double out_vec[6], double out_cov[36], double* path,
                     uint32_t* n_steps) {
    using namespace oracle;
    bound_state<double> in, o;
    in.surface = start;
    for (int i = 0; i < 6; ++i) in.vec.a[i][0] = vec[i];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) in.cov.a[r][c] = cov[6 * r + c];
    const double mass = cfg->ptc_mass, aq = std::fabs(cfg->ptc_charge);
    if (apply_start_material) add_material<double>(det->surfaces[start], in, mass, aq);
    const int r = propagate_bound<double>(*cfg, *det, *field, in, mass, aq, o, path, n_steps);
    if (r == 0) {
        add_material<double>(det->surfaces[o.surface], o, mass, aq);
        *end_surface = o.surface;
        for (int i = 0; i < 6; ++i) out_vec[i] = o.vec.a[i][0];
        for (int rr = 0; rr < 6; ++rr)
            for (int c = 0; c < 6; ++c) out_cov[6 * rr + c] = o.cov.a[rr][c];
    }
    return r;
}

int oracle_propagate_direct(const ckf_finding_config* cfg, const ckf_detector* det,
                            const ckf_bfield* field, uint32_t start, uint32_t target,
                            const double vec[6], const double cov[36], double out_vec[6],
                            double out_cov[36]) {
    using namespace oracle;
    bound_state<double> in, o;
    in.surface = start;
    for (int i = 0; i < 6; ++i) in.vec.a[i][0] = vec[i];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) in.cov.a[r][c] = cov[6 * r + c];
    const int r = propagate_direct<double>(*cfg, *det, *field, in, target, cfg->ptc_mass,
                                           std::fabs(cfg->ptc_charge), o);
    if (r == 0) {
        for (int i = 0; i < 6; ++i) out_vec[i] = o.vec.a[i][0];
        for (int rr = 0; rr < 6; ++rr)
            for (int c = 0; c < 6; ++c) out_cov[6 * rr + c] = o.cov.a[rr][c];
    }
    return r;
}

void oracle_rk_step(const ckf_bfield* field, double mass, double abs_q, double h, const double free_in[8],
                    double free_out[8], double D[64]) {
    using namespace oracle;
    mat<double, 8, 1> F;
    for (int i = 0; i < 8; ++i) F.a[i][0] = free_in[i];
    mat<double, 8, 8> M = mat<double, 8, 8>::identity();
    rk_step<double>(*field, mass, abs_q, h, F, &M);
    for (int i = 0; i < 8; ++i) free_out[i] = F.a[i][0];
    for (int r = 0; r < 8; ++r)
        for (int c = 0; c < 8; ++c) D[8 * r + c] = M.a[r][c];
}

}  // extern "C"
