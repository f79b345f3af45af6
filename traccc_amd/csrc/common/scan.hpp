// Device-wide exclusive scan of n u32 values into out[0 .. n] (out[n] = the
// total), in three passes: per-tile scans with the tile totals, one
// workgroup scanning the totals, then the tile offsets added.  Each
// translation unit that includes this header gets its own kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../ckf/wave.hpp"
#include "errors.hpp"

namespace tamd {
namespace {

constexpr uint32_t kScanBlock = 256;
constexpr uint32_t kScanTopBlock = 1024;
constexpr uint32_t kScanPer = 4;                       // items per thread
constexpr uint32_t kScanTile = kScanBlock * kScanPer;  // items per tile

__global__ __launch_bounds__(kScanBlock) void k_scan_local(const uint32_t* in, uint32_t* out,
                                                           uint32_t* sums, uint32_t n) {
    __shared__ uint32_t s_w[kScanBlock / kWave];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    uint32_t v[kScanPer], sum = 0u;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        sum += v[k];
    }
    const uint32_t incl = wave_inclusive_scan(sum);
    const uint32_t w = threadIdx.x / kWave;
    if (lane() == kWave - 1) s_w[w] = incl;
    __syncthreads();
    uint32_t woff = 0u;
    for (uint32_t q = 0; q < w; ++q) woff += s_w[q];
    uint32_t ex = woff + incl - sum;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == blockDim.x - 1u) sums[blockIdx.x] = woff + incl;
}

__global__ __launch_bounds__(kScanTopBlock) void k_scan_sums(uint32_t* sums, uint32_t nb, uint32_t* total,
                                                             uint32_t* total_copy) {
    __shared__ uint32_t s_w[kScanTopBlock / kWave];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0u) carry = 0u;
    __syncthreads();
    const uint32_t w = threadIdx.x / kWave;
    for (uint32_t c0 = 0; c0 < nb; c0 += kScanTopBlock) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        const uint32_t incl = wave_inclusive_scan(v);
        if (lane() == kWave - 1) s_w[w] = incl;
        __syncthreads();
        uint32_t woff = 0u;
        for (uint32_t q = 0; q < w; ++q) woff += s_w[q];
        const uint32_t ex = carry + woff + incl - v;
        if (i < nb) sums[i] = ex;
        __syncthreads();
        if (threadIdx.x == kScanTopBlock - 1u) carry = ex + v;
        __syncthreads();
    }
    if (threadIdx.x == 0u) {
        *total = carry;
        if (total_copy != nullptr) *total_copy = carry;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(uint32_t* out, const uint32_t* sums, uint32_t n) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    const uint32_t add = sums[blockIdx.x];
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k) {
        if (base + k < n) out[base + k] += add;
    }
}

// Tile totals needed by exclusive_scan for n items.
inline uint32_t scan_sums_size(uint32_t n) { return (n + kScanTile - 1u) / kScanTile + 1u; }

// out[0 .. n] = exclusive scan of in[0 .. n); total_copy (device, may be
// null) also receives the total.  sums: scan_sums_size(n) u32 of scratch.
inline void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t* sums, uint32_t n,
                           uint32_t* total_copy, hipStream_t s) {
    const uint32_t nb = (n + kScanTile - 1u) / kScanTile;
    if (nb > 0u) {
        hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kScanBlock), 0, s, in, out, sums, n);
    }
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanTopBlock), 0, s, sums, nb, out + n, total_copy);
    if (nb > 0u) {
        hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanBlock), 0, s, out, sums, n);
    }
    hip_check(hipGetLastError(), "scan");
}

}  // namespace
}  // namespace tamd
