// Measurement-vote pruning of the track candidates (traccc device chain,
// max_num_tracks_per_measurement > 0):
//   gather_best_tips_per_measurement  device/common/include/traccc/finding/device/impl/gather_best_tips_per_measurement.ipp:27-180
//   gather_measurement_votes          device/common/include/traccc/finding/device/impl/gather_measurement_votes.ipp:19-47
//   update_tip_length_buffer          device/common/include/traccc/finding/device/impl/update_tip_length_buffer.ipp:20-54
//   orchestration                     device/common/src/finding/combinatorial_kalman_filter_algorithm.cpp:568-634
//
// Every measurement keeps the K tracks with the highest p-value of their
// tip's running (chi2, ndf) sums; a track survives when the measurements that
// kept it make up at least min_measurement_voting_fraction of its
// measurements.
//
// The reference serialises each measurement's top-K list behind a 64-bit
// insertion mutex.  Here every (track, measurement) pair instead runs a
// lock-free insertion cascade of 64-bit atomicMax operations down the
// measurement's K slots, on a key (p-value bits, ~track index): the larger key
// stays, the smaller one moves on, and whatever leaves slot K-1 is out.
// Each key is at exactly one place (a slot or its carrier) at any time and
// slot values only grow, so after every cascade has ended the slots hold
// exactly the K largest keys, whatever the interleaving.  Equal p-values
// rank by lower track index.  Survivors keep their order (a scan with
// decoupled look-back), so the result is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_types.hpp"
#include "kalman_device.hpp"
#include "kernels.hpp"
#include "wave.hpp"

namespace tamd {

namespace {

constexpr int kBlock = 256;

__device__ inline uint32_t event_of_track(const uint32_t* track_off, uint32_t n_events, uint32_t t) {
    // last e with track_off[e] <= t
    uint32_t lo = 0u, hi = n_events;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) / 2u;
        if (track_off[mid] <= t) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void k_vote_insert(vote_params V) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < V.n_tracks;
         t += gridDim.x * blockDim.x) {
        const uint32_t ev = event_of_track(V.track_off, V.n_events, t);
        const uint32_t moff = V.event_meas_off[ev];
        const uint64_t key0 = V.key[t];
        for (uint32_t c = V.offsets[t]; c < V.offsets[t + 1u]; ++c) {
            unsigned long long* slot =
                reinterpret_cast<unsigned long long*>(V.slots) + static_cast<size_t>(moff + V.cons[c]) * V.K;
            unsigned long long key = key0;
            for (uint32_t k = 0; k < V.K; ++k) {
                const unsigned long long old = atomicMax(&slot[k], key);
                if (old < key) {
                    if (old == 0ull) {
                        break;  // took an empty slot: nothing displaced
                    }
                    key = old;  // carry the displaced key down
                }
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_vote_count(vote_params V) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < V.n_tracks;
         t += gridDim.x * blockDim.x) {
        const uint32_t ev = event_of_track(V.track_off, V.n_events, t);
        const uint32_t moff = V.event_meas_off[ev];
        const uint64_t key = V.key[t];
        const uint32_t c0 = V.offsets[t], c1 = V.offsets[t + 1u];
        uint32_t votes = 0u;
        for (uint32_t c = c0; c < c1; ++c) {
            const uint64_t* slot = V.slots + static_cast<size_t>(moff + V.cons[c]) * V.K;
            for (uint32_t k = 0; k < V.K; ++k) {
                if (slot[k] == key) {
                    ++votes;
                    break;
                }
            }
        }
        const uint32_t n = c1 - c0;
        // update_tip_length_buffer.ipp:36-42: float fraction, dropped below the minimum
        const float frac = static_cast<float>(votes) / static_cast<float>(n);
        V.keep[t] = (frac < V.min_fraction) ? 0u : 1u;
    }
}

// Order-preserving compaction: per track the kept rank and constituent
// position; per event the new offsets.
__global__ __launch_bounds__(kBlock) void k_vote_scan(vote_params V) {
    step_ctl* ctl = V.ctl;
    const uint32_t n = V.n_tracks;
    const uint32_t epoch = ctl->epoch;
    const uint32_t n_parts = (n + kWave - 1) / kWave;
    for (;;) {
        const uint32_t p = wave_ticket(&ctl->ticket[1].v);
        if (p >= n_parts) {
            break;
        }
        const uint32_t t = p * kWave + lane();
        const bool valid = t < n;
        uint32_t keep = 0u, nc = 0u;
        if (valid) {
            keep = V.keep[t];
            nc = V.offsets[t + 1u] - V.offsets[t];
        }
        const uint32_t w = keep ? nc : 0u;
        const uint32_t cincl = wave_inclusive_scan(w);
        const uint64_t mk = wave_ballot(keep != 0u);
        const uint32_t k_tot = static_cast<uint32_t>(__popcll(mk));
        const uint32_t c_tot = wave_broadcast(cincl, kWave - 1);
        const uint64_t b = wave_lookback(V.lb_lo, V.lb_hi, p, epoch, pack2(k_tot, c_tot));
        if (valid) {
            V.rank[t] = lo32(b) + prefix_count(mk);
            V.cpos[t] = hi32(b) + cincl - w;
        }
        if (p == n_parts - 1u && lane() == 0u) {
            V.totals[0] = lo32(b) + k_tot;
            V.totals[1] = hi32(b) + c_tot;
        }
    }
}

__global__ void k_vote_events(vote_params V) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e <= V.n_events;
         e += gridDim.x * blockDim.x) {
        const uint32_t t = V.track_off[e];
        if (t < V.n_tracks) {
            V.new_track_off[e] = V.rank[t];
            V.new_const_off[e] = V.cpos[t];
        } else {
            V.new_track_off[e] = V.totals[0];
            V.new_const_off[e] = V.totals[1];
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_vote_write(vote_params V) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < V.n_tracks;
         t += gridDim.x * blockDim.x) {
        if (V.keep[t] == 0u) {
            continue;
        }
        const uint32_t r = V.rank[t];
        const uint32_t cp = V.cpos[t];
        const uint32_t c0 = V.offsets[t], c1 = V.offsets[t + 1u];
        for (uint32_t c = c0; c < c1; ++c) {
            V.o_cons[cp + (c - c0)] = V.cons[c];
        }
        V.o_offsets[r] = cp;
        V.o_outcome[r] = V.outcome[t];
        V.o_surface[r] = V.surface[t];
        copy_track_params(V.vec + 6u * static_cast<size_t>(t), V.cov + 36u * static_cast<size_t>(t),
                          V.o_vec + 6u * static_cast<size_t>(r), V.o_cov + 36u * static_cast<size_t>(r));
        V.o_ndf[r] = V.ndf[t];
        V.o_chi2[r] = V.chi2[t];
        V.o_pval[r] = V.pval[t];
        V.o_nholes[r] = V.nholes[t];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        V.o_offsets[V.totals[0]] = V.totals[1];
    }
}

}  // namespace

hipError_t launch_votes(const vote_params& V, uint32_t n_cu, hipStream_t stream) {
    const uint32_t blocks = std::max(1u, std::min((V.n_tracks + kBlock - 1u) / kBlock, n_cu * 8u));
    hipLaunchKernelGGL(k_vote_insert, dim3(blocks), dim3(kBlock), 0, stream, V);
    hipLaunchKernelGGL(k_vote_count, dim3(blocks), dim3(kBlock), 0, stream, V);
    const uint32_t parts = (V.n_tracks + kWave - 1u) / kWave;
    const uint32_t scan_blocks = std::max(1u, std::min((parts + 3u) / 4u, n_cu * 8u));
    hipLaunchKernelGGL(k_vote_scan, dim3(scan_blocks), dim3(kBlock), 0, stream, V);
    hipLaunchKernelGGL(k_vote_events, dim3((V.n_events + 1u + 255u) / 256u), dim3(256), 0, stream, V);
    hipLaunchKernelGGL(k_vote_write, dim3(blocks), dim3(kBlock), 0, stream, V);
    return hipGetLastError();
}

}  // namespace tamd
