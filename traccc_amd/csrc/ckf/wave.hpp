// Wavefront (64-lane) primitives for branch compaction on CDNA:
//  * ballot/mbcnt prefix counts for 0/1 flags,
//  * DPP-free shuffle scans for small integer counts,
//  * a per-wave work queue (one atomic ticket per 64 items),
//  * decoupled look-back across waves (Merrill & Garland, "Single-pass
//    Parallel Prefix Scan with Decoupled Look-back", NVR-2016-002), with
//    separate aggregate / inclusive slots and an epoch tag so descriptor
//    arrays are reused by every step without clearing.
// Waves never synchronise with each other except through the look-back,
// whose predecessor chain always ends at partition 0; tickets are handed out
// in increasing order, so every predecessor belongs to a running wave.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tamd {

constexpr int kWave = 64;

__device__ inline uint32_t lane() { return __lane_id(); }

__device__ inline uint32_t prefix_count(uint64_t mask) {
    // number of set bits of `mask` in lanes below this one
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ inline uint64_t wave_ballot(bool p) { return __ballot(p); }

// Inclusive wave scan of a u32.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t l = lane();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, kWave);
        if (l >= static_cast<uint32_t>(d)) {
            v += o;
        }
    }
    return v;
}

__device__ inline uint32_t wave_broadcast(uint32_t v, int src) { return __shfl(v, src, kWave); }

// Waves of a ticketed grid that join the queue.  A grid sized for the
// largest step holds many more waves than a small step has partitions; waves
// past the first n_parts take no ticket (each would only fail one more atomic
// on the shared counter).  Wave 0 always joins, so work after the queue
// still has a wave to run it.  Partition p is only waited on by tickets after
// it, which running waves hold, so the queue drains with any set of joiners.
__device__ inline uint32_t ticket_waves(uint32_t n_parts) {
    const uint32_t all = gridDim.x * (blockDim.x / kWave);
    const uint32_t want = n_parts > 0u ? n_parts : 1u;
    return want < all ? want : all;
}
__device__ inline bool wave_takes_tickets(uint32_t n_parts) {
    return blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave < ticket_waves(n_parts);
}

// One track's 6 parameters and 6 x 6 covariance (24 B and 144 B rows, so
// 8 / 16 B pieces stay aligned): 3 + 9 memory instructions instead of 42
// 4-byte accesses strided across the wave.  src_cov == nullptr writes zeros.
__device__ inline void copy_track_params(const float* src_vec, const float* src_cov,
                                         float* dst_vec, float* dst_cov) {
    const float2* sv = reinterpret_cast<const float2*>(src_vec);
    const float4* sc = reinterpret_cast<const float4*>(src_cov);
    float2* dv = reinterpret_cast<float2*>(dst_vec);
    float4* dc = reinterpret_cast<float4*>(dst_cov);
#pragma unroll
    for (int k = 0; k < 3; ++k) dv[k] = src_vec != nullptr ? sv[k] : make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 9; ++k) dc[k] = src_cov != nullptr ? sc[k] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Next work partition for the calling wave.
__device__ inline uint32_t wave_ticket(uint32_t* counter) {
    uint32_t t = 0;
    if (lane() == 0) {
        t = atomicAdd(counter, 1u);
    }
    return wave_broadcast(t, 0);
}

// Partitions handed out several per ticket: consecutive partitions (the
// branches of neighbouring seeds, which read the same surfaces) stay on one
// wave, and a queue of millions of small partitions takes a quarter of the
// atomics on its counters.  Small queues keep one partition per ticket so
// every wave still gets work.  next() returns the wave's next partition
// (n_parts when the queue is empty).
//
// The tickets form kXcds contiguous ranges with a counter each (128 B
// apart); a wave takes from its own XCD's range first -- neighbouring
// partitions read the same measurement lines, which then stay in that XCD's
// L2 -- and moves on to the next range when its own is spent, until every
// range is.  Workgroups are dispatched to the XCDs round-robin, so block b
// runs on XCD b % kXcds (a different placement only costs locality).  A
// queue of fewer than kXcdMinTickets tickets stays one range: its waves
// would hold a ticket or two each and then spend more time failing on the
// other ranges' counters than the locality saves.  Only
// for kernels without look-back: a partition would otherwise wait on every
// partition of the lower ranges.
constexpr uint32_t kXcds = 8u;
constexpr uint32_t kXcdCounterStride = 32u;  // u32 words between counters
constexpr uint32_t kXcdMinTickets = 1024u;

struct xcd_ticket_run {
    uint32_t* counters;
    uint32_t n_parts;
    uint32_t per;
    uint32_t n_tickets;
    bool takes;
    uint32_t n_ranges;    // kXcds, or 1 for a small queue
    uint32_t home;        // this block's range: its XCD
    uint32_t tried = 0u;  // ranges found spent, in order from home
    uint32_t base = 0u;
    uint32_t sub = 0u;
    __device__ xcd_ticket_run(uint32_t* c, uint32_t n, uint32_t max_per = 4u)
        : counters(c), n_parts(n), per(per_ticket(n) < max_per ? per_ticket(n) : max_per) {
        n_tickets = (n_parts + per - 1u) / per;
        takes = wave_takes_tickets(n_tickets);
        n_ranges = n_tickets >= kXcdMinTickets ? kXcds : 1u;
        home = blockIdx.x % n_ranges;
        sub = per;  // no ticket held
    }
    // one more partition per ticket for every 4,096 partitions, up to four
    __device__ static uint32_t per_ticket(uint32_t n) {
        const uint32_t q = n / 4096u;
        return q < 1u ? 1u : (q > 4u ? 4u : q);
    }
    __device__ uint32_t next() {
        if (sub == per) {
            uint32_t t = n_tickets;
            while (takes && tried < n_ranges) {
                const uint32_t x = (home + tried) % n_ranges;
                const uint32_t lo = static_cast<uint32_t>(uint64_t(n_tickets) * x / n_ranges);
                const uint32_t hi = static_cast<uint32_t>(uint64_t(n_tickets) * (x + 1u) / n_ranges);
                if (lo < hi) {
                    const uint32_t got = lo + wave_ticket(counters + x * kXcdCounterStride);
                    if (got < hi) {
                        t = got;
                        break;
                    }
                }
                ++tried;
            }
            base = t * per;
            sub = 0u;
        }
        const uint32_t p = base + sub;
        ++sub;
        return p < n_parts ? p : n_parts;
    }
};

__device__ inline uint64_t ld_relaxed64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Two 32-bit counters packed in one u64 (sums never carry across halves).
__device__ inline uint64_t pack2(uint32_t lo, uint32_t hi) {
    return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}
__device__ inline uint32_t lo32(uint64_t v) { return static_cast<uint32_t>(v); }
__device__ inline uint32_t hi32(uint64_t v) { return static_cast<uint32_t>(v >> 32); }

// Decoupled look-back descriptors: one 64-bit word per partition and
// channel holding the value with its status, so a single relaxed device-scope
// atomic carries both and no release / acquire fence is needed (on this GPU
// such fences write back / invalidate the whole L2 of the XCD, which made
// look-back kernels spend most of their time in cache maintenance):
//   value (32 b) << 32 | (epoch mod 2^30) << 2 | status (1 aggregate,
//   2 inclusive prefix; any other epoch reads as not ready).
__device__ inline uint64_t lb_word(uint32_t value, uint32_t epoch, uint32_t status) {
    return (static_cast<uint64_t>(value) << 32) |
           static_cast<uint64_t>(((epoch & 0x3fffffffu) << 2) | status);
}

// Exclusive prefix of one channel over partitions 0..p-1 (the aggregate of
// p already published).  The whole wave inspects the 64 preceding words per
// round trip (lane l reads partition q - 1 - l), stops at the nearest
// inclusive prefix, and sums the aggregates in between.
__device__ inline uint32_t lookback_channel(uint64_t* words, uint32_t p, uint32_t epoch,
                                            uint32_t agg) {
    const uint32_t l = lane();
    const uint32_t tag = epoch & 0x3fffffffu;
    uint32_t excl = 0u;
    int64_t window_end = static_cast<int64_t>(p);  // partitions below this are unread
    for (;;) {
        const int64_t q = window_end - 1 - static_cast<int64_t>(l);
        uint32_t status = 2u;  // lanes beyond partition 0 act as a terminating prefix of 0
        uint32_t v = 0u;
        if (q >= 0) {
            const uint64_t w = ld_relaxed64(&words[q]);
            status = ((lo32(w) >> 2) == tag) ? (lo32(w) & 3u) : 0u;
            v = hi32(w);
        }
        const uint64_t m_pref = wave_ballot(status == 2u);
        const uint64_t m_ready = wave_ballot(status == 1u || status == 2u);
        // lanes up to and including the nearest prefix must all be ready
        const uint32_t first_pref = m_pref ? static_cast<uint32_t>(__ffsll(static_cast<long long>(m_pref)) - 1) : 64u;
        const uint64_t need = (first_pref >= 63u) ? ~0ull : ((2ull << first_pref) - 1ull);
        if ((m_ready & need) != need) {
            __builtin_amdgcn_s_sleep(1);
            continue;  // re-read the same window
        }
        uint32_t take = (l <= first_pref) ? v : 0u;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) take += __shfl_xor(take, d, kWave);
        excl += take;
        if (first_pref < 64u) {
            break;
        }
        window_end -= kWave;
    }
    if (l == 0u) {
        st_relaxed64(&words[p], lb_word(excl + agg, epoch, 2u));
    }
    return excl;
}

// Inclusive prefix of partition q once published (status 2 of this epoch);
// waits for it.  Used after every partition's wave has left the queue.
__device__ inline uint32_t lookback_inclusive(const uint64_t* words, uint32_t q, uint32_t epoch) {
    const uint32_t want = ((epoch & 0x3fffffffu) << 2) | 2u;
    for (;;) {
        const uint64_t w = ld_relaxed64(&words[q]);
        if (lo32(w) == want) {
            return hi32(w);
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Exclusive prefix of partition p's packed pair of counters (low, high) over
// partitions 0..p-1.  Called by all lanes of the wave.  hi_words may be null
// when the high counter is always 0.  Both aggregates are published before
// either channel is looked back on, so no partition waits on a channel its
// predecessor has not published.
__device__ inline uint64_t wave_lookback(uint64_t* lo_words, uint64_t* hi_words, uint32_t p,
                                         uint32_t epoch, uint64_t agg) {
    const uint32_t status = p == 0u ? 2u : 1u;
    if (lane() == 0u) {
        st_relaxed64(&lo_words[p], lb_word(lo32(agg), epoch, status));
        if (hi_words != nullptr) {
            st_relaxed64(&hi_words[p], lb_word(hi32(agg), epoch, status));
        }
    }
    if (p == 0u) {
        return 0ull;
    }
    const uint32_t lo = lookback_channel(lo_words, p, epoch, lo32(agg));
    const uint32_t hi = hi_words != nullptr ? lookback_channel(hi_words, p, epoch, hi32(agg)) : 0u;
    return pack2(lo, hi);
}

}  // namespace tamd
