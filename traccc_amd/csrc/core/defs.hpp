// Common definitions shared by the HIP kernels and the host-side simulator.
//
// Units follow traccc/detray (core/include/traccc/definitions/common.hpp:18-25):
// mm, GeV, ns, elementary charge.  A field in tesla is converted to the
// internal unit GeV/(e*mm) by multiplying with unit_T.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TAMD_HD __host__ __device__ inline
#define TAMD_DEV __device__ inline
// full unrolling keeps small matrices in registers (a runtime index into a
// register array sends it to scratch memory)
#define TAMD_UNROLL _Pragma("unroll")
#define TAMD_NO_UNROLL _Pragma("unroll 1")
#else
#define TAMD_HD inline
#define TAMD_DEV inline
#define TAMD_UNROLL
#define TAMD_NO_UNROLL
#endif

#include "traccc_amd/ckf.h"

namespace tamd {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double unit_T = 0.000299792458;     // tesla -> GeV / (e mm)
constexpr double kSpeedOfLight = 299.792458;  // mm / ns

// Bound parameter indices (traccc/detray e_bound_*):
// loc0, loc1, phi, theta, q/p, time.
enum bound_index : int {
    e_loc0 = 0,
    e_loc1 = 1,
    e_phi = 2,
    e_theta = 3,
    e_qop = 4,
    e_time = 5
};

// Packed symmetric 6x6 index (row-major upper triangle).
TAMD_HD constexpr int sym_idx(int i, int j) {
    // i <= j assumed by callers that use sym6; general form below
    return (i <= j) ? (i * 6 - (i * (i - 1)) / 2 + (j - i))
                    : (j * 6 - (j * (j - 1)) / 2 + (i - j));
}

template <typename T>
struct vec3 {
    T x, y, z;
};

template <typename T>
TAMD_HD vec3<T> operator+(vec3<T> a, vec3<T> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename T>
TAMD_HD vec3<T> operator-(vec3<T> a, vec3<T> b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename T>
TAMD_HD vec3<T> operator*(T s, vec3<T> a) {
    return {s * a.x, s * a.y, s * a.z};
}
template <typename T>
TAMD_HD T dot(vec3<T> a, vec3<T> b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T>
TAMD_HD vec3<T> cross(vec3<T> a, vec3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
            a.x * b.y - a.y * b.x};
}

}  // namespace tamd
