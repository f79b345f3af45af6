// Magnetic field lookup: constant vector or an (r, z) table with bilinear
// interpolation.  Replaces the covfie backends traccc instantiates
// (core/include/traccc/bfield/magnetic_field_types.hpp:22-55): the constant
// field and an inhomogeneous field on a regular grid.
#pragma once

#include <cmath>

#include "defs.hpp"

namespace tamd {

template <typename T>
struct field_view {
    uint32_t kind;      // ckf_bfield_kind
    T b[3];             // constant field, internal units
    const float* map;   // [n_z][n_r][2] (B_r, B_z) in tesla
    int n_r, n_z;
    T r_min, z_min, inv_dr, inv_dz;
};

template <typename T>
inline field_view<T> make_field_view(const ckf_bfield& f, const float* map_ptr) {
    field_view<T> v{};
    v.kind = f.kind;
    for (int i = 0; i < 3; ++i) {
        v.b[i] = static_cast<T>(f.b[i] * unit_T);
    }
    v.map = map_ptr;
    v.n_r = static_cast<int>(f.n_r);
    v.n_z = static_cast<int>(f.n_z);
    v.r_min = static_cast<T>(f.r_min);
    v.z_min = static_cast<T>(f.z_min);
    v.inv_dr = f.dr > 0.f ? static_cast<T>(1.0 / f.dr) : T(0);
    v.inv_dz = f.dz > 0.f ? static_cast<T>(1.0 / f.dz) : T(0);
    return v;
}

// Field at a global position, internal units.  The table is clamped at its
// edges (a point outside the grid sees the nearest edge cell).
template <typename T>
TAMD_HD vec3<T> field_at(const field_view<T>& f, const vec3<T>& p) {
    if (f.kind == CKF_BFIELD_CONSTANT) {
        return {f.b[0], f.b[1], f.b[2]};
    }
    const T r = std::sqrt(p.x * p.x + p.y * p.y);
    T fr = (r - f.r_min) * f.inv_dr;
    T fz = (p.z - f.z_min) * f.inv_dz;
    const T max_r = static_cast<T>(f.n_r - 1);
    const T max_z = static_cast<T>(f.n_z - 1);
    fr = fr < T(0) ? T(0) : (fr > max_r ? max_r : fr);
    fz = fz < T(0) ? T(0) : (fz > max_z ? max_z : fz);
    int ir = static_cast<int>(fr);
    int iz = static_cast<int>(fz);
    ir = ir > f.n_r - 2 ? f.n_r - 2 : ir;
    iz = iz > f.n_z - 2 ? f.n_z - 2 : iz;
    const T tr = fr - static_cast<T>(ir);
    const T tz = fz - static_cast<T>(iz);
    const float* c00 = f.map + 2 * (iz * f.n_r + ir);
    const float* c01 = c00 + 2;              // ir + 1
    const float* c10 = c00 + 2 * f.n_r;      // iz + 1
    const float* c11 = c10 + 2;
    const T w00 = (T(1) - tr) * (T(1) - tz);
    const T w01 = tr * (T(1) - tz);
    const T w10 = (T(1) - tr) * tz;
    const T w11 = tr * tz;
    const T br = (w00 * c00[0] + w01 * c01[0] + w10 * c10[0] + w11 * c11[0]) *
                 static_cast<T>(unit_T);
    const T bz = (w00 * c00[1] + w01 * c01[1] + w10 * c10[1] + w11 * c11[1]) *
                 static_cast<T>(unit_T);
    if (r > T(0)) {
        const T inv_r = T(1) / r;
        return {br * p.x * inv_r, br * p.y * inv_r, bz};
    }
    return {T(0), T(0), bz};
}

}  // namespace tamd
