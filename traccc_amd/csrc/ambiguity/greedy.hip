// Greedy ambiguity resolution for MI355X (C ABI: include/traccc_amd/ambiguity.h).
//
// The reference (core/src/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.cpp:19-254)
// removes one track per iteration: the worst accepted track by
// (rel_shared = n_shared / n_meas descending, p-value ascending), where
// n_shared counts the track's constituents whose measurement identifier is
// held by another accepted track; a removal that leaves one track on an
// identifier takes that identifier out of the survivor's n_shared.  Its CUDA
// port keeps a globally sorted order and removes a run of tracks from its
// tail per graph iteration.
//
// Here the same result comes from rounds with no global order.  Call a
// track "best on" an identifier when it is the best (least-worse) accepted
// track among the tracks holding it.  With max_shared_meas = 1, every
// accepted track t with n_shared >= 1 that is best on none of its shared
// identifiers is removed by the sequential algorithm, with its current
// n_shared: each shared identifier of t keeps its best track, which is
// better than t and so cannot be removed while t is accepted; t's n_shared
// therefore stays, the tracks removed before t are worse than t (and keep
// n_shared >= 1, so the iteration does not stop first), and removing t
// early changes no decision of the sequential run.  Such tracks stay in that
// state when other such tracks go, so one round removes all of them:
//
//   k_amb_best    per shared identifier (at least two accepted tracks): its
//                 best track; that track is marked protected for the round
//   k_amb_remove  per candidate track (accepted, n_shared >= 1): not
//                 protected -> removed; for each of its distinct identifiers
//                 the accepted count drops, and the removal that leaves one
//                 track there (the protected best) lowers that track's
//                 n_shared by the identifier's multiplicity in it
//
// The global worst track is always removable, so every round removes at
// least one track; a round that removes none ends the run (then no track
// shares a measurement).  The rounds remove exactly the sequential
// algorithm's tracks; their number is the sequential iteration count.
// max_shared_meas > 1 (the stop test is then global) and a max_iterations
// below that count (the sequential algorithm would stop part-way) run the
// one-at-a-time path instead: one workgroup that finds the worst track and
// the largest n_shared, removes the track, and repeats.
//
// Identifiers index per-identifier arrays (counts, track lists, bests) of
// id_bound entries; the tracks of each identifier are a CSR built by a
// count, an exclusive scan and a scatter.  Lists of the still-shared
// identifiers and of the candidate tracks shrink round by round.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "../ckf/handles.hpp"
#include "../ckf/wave.hpp"
#include "../common/errors.hpp"
#include "../common/scan.hpp"
#include "traccc_amd/ambiguity.h"

namespace tamd {
namespace {

constexpr uint32_t kAmbBlock = 256;
constexpr uint32_t kSeqBlock = 1024;
constexpr uint32_t kDead = 0u, kAlive = 1u;
constexpr uint32_t kNone = 0xffffffffu;

struct amb_ctl {
    uint32_t n_hot[2];      // shared-identifier lists, by round parity
    uint32_t n_cand[2];     // candidate-track lists, by round parity
    uint32_t removed[2];    // removals of a round, by round parity
    uint32_t done;
    uint32_t last_round;    // the round that found the previous one empty
    uint32_t total_removed;
    uint32_t bad_input;     // an identifier out of range or a decreasing offset
    uint32_t n_candidates;  // tracks with at least min_meas_per_track measurements
    uint32_t n_accepted;
    uint32_t scan_total;    // distinct (identifier, track) pairs
    uint32_t pad[5];
};

struct amb_params {
    uint32_t n_tracks;
    uint32_t id_base;   // identifiers are in [id_base, id_base + id_bound)
    uint32_t id_bound;
    uint32_t min_meas;
    uint32_t n_entries;    // constituents: off[n_tracks] of a valid input
    const uint32_t* off;   // [n_tracks + 1]
    const uint32_t* ids;   // absolute constituent positions off[t] ..
    const float* pval;
    uint32_t* status;      // [n_tracks]
    uint32_t* n_shared;    // [n_tracks]
    uint32_t* protect;     // [n_tracks] round tag (r + 1) of the tracks best on some identifier
    uint32_t* cand[2];     // [n_tracks]
    uint32_t* n_acc;       // [id_bound] accepted tracks holding the identifier
    uint32_t* list_off;    // [id_bound + 1]
    uint32_t* fill;        // [id_bound]
    uint32_t* list;        // [pairs] the tracks of each identifier
    uint32_t* best_of;     // [id_bound]
    uint32_t* hot[2];      // [id_bound]
    amb_ctl* ctl;
};

// A track's place in the order: rel_shared, then p-value (NaN lowest), then
// the index.  worse(a, b): a is removed before b.
struct track_key {
    float rel;
    float pk;
    uint32_t t;
};

__device__ inline track_key key_of(const amb_params& P, uint32_t t) {
    const uint32_t nm = P.off[t + 1] - P.off[t];
    const uint32_t ns = P.n_shared[t];
    track_key k;
    k.rel = nm == 0u ? 0.f : __fdiv_rn(static_cast<float>(ns), static_cast<float>(nm));
    const float p = P.pval[t];
    k.pk = isnan(p) ? -INFINITY : p;
    k.t = t;
    return k;
}

__device__ inline bool worse(const track_key& a, const track_key& b) {
    if (a.rel != b.rel) return a.rel > b.rel;
    if (a.pk != b.pk) return a.pk < b.pk;
    return a.t > b.t;
}

// Is constituent j the first of its identifier in the track's list [b, j]?
__device__ inline bool first_occurrence(const uint32_t* ids, uint32_t b, uint32_t j) {
    const uint32_t id = ids[j];
    for (uint32_t q = b; q < j; ++q) {
        if (ids[q] == id) return false;
    }
    return true;
}

__device__ inline uint32_t multiplicity(const amb_params& P, uint32_t t, uint32_t id) {
    uint32_t c = 0u;
    for (uint32_t q = P.off[t]; q < P.off[t + 1]; ++q) c += P.ids[q] == id ? 1u : 0u;
    return c;
}

// Append `value` for the lanes with `take` (called by the whole wave).
__device__ inline void wave_append(uint32_t* counter, uint32_t* list, bool take, uint32_t value) {
    const uint64_t m = wave_ballot(take);
    if (m == 0ull) return;
    const int leader = __builtin_ctzll(m);
    uint32_t base = 0u;
    if (lane() == static_cast<uint32_t>(leader)) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, kWave);
    if (take) list[base + prefix_count(m)] = value;
}

__device__ inline void wave_count(uint32_t* counter, bool p) {
    const uint64_t m = wave_ballot(p);
    if (m != 0ull && lane() == static_cast<uint32_t>(__builtin_ctzll(m))) {
        atomicAdd(counter, __popcll(m));
    }
}

// Offsets non-decreasing and every identifier in range; a bad input leaves
// the flag and takes no part in the counts (no access out of bounds).
__global__ __launch_bounds__(kAmbBlock) void k_amb_count(amb_params P) {
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < P.n_tracks; b0 += gridDim.x * blockDim.x) {
        const uint32_t t = b0 + threadIdx.x;
        bool cand = false;
        if (t < P.n_tracks) {
            const uint32_t b = P.off[t], e = P.off[t + 1];
            P.n_shared[t] = 0u;
            P.protect[t] = 0u;
            if (e < b || e > P.n_entries) {
                atomicOr(&P.ctl->bad_input, 1u);
                P.status[t] = kDead;
            } else {
                cand = e - b >= P.min_meas;
                P.status[t] = cand ? kAlive : kDead;
                for (uint32_t j = b; cand && j < e; ++j) {
                    const uint32_t id = P.ids[j] - P.id_base;
                    if (P.ids[j] < P.id_base || id >= P.id_bound) {
                        atomicOr(&P.ctl->bad_input, 2u);
                        continue;
                    }
                    if (first_occurrence(P.ids, b, j)) atomicAdd(&P.n_acc[id], 1u);
                }
            }
        }
        wave_count(&P.ctl->n_candidates, cand);
    }
}

// The identifiers' track lists, each track's n_shared, and the first
// candidate list (accepted tracks with n_shared >= 1).
__global__ __launch_bounds__(kAmbBlock) void k_amb_fill(amb_params P) {
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < P.n_tracks; b0 += gridDim.x * blockDim.x) {
        const uint32_t t = b0 + threadIdx.x;
        bool shared = false;
        if (t < P.n_tracks && P.status[t] == kAlive) {
            const uint32_t b = P.off[t], e = P.off[t + 1];
            uint32_t ns = 0u;
            for (uint32_t j = b; j < e; ++j) {
                const uint32_t id = P.ids[j] - P.id_base;
                if (P.ids[j] < P.id_base || id >= P.id_bound) continue;
                if (first_occurrence(P.ids, b, j)) {
                    P.list[P.list_off[id] + atomicAdd(&P.fill[id], 1u)] = t;
                }
                ns += P.n_acc[id] >= 2u ? 1u : 0u;
            }
            P.n_shared[t] = ns;
            shared = ns >= 1u;
        }
        wave_append(&P.ctl->n_cand[0], P.cand[0], shared, t);
    }
}

__global__ __launch_bounds__(kAmbBlock) void k_amb_hot0(amb_params P) {
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < P.id_bound; b0 += gridDim.x * blockDim.x) {
        const uint32_t m = b0 + threadIdx.x;
        wave_append(&P.ctl->n_hot[0], P.hot[0], m < P.id_bound && P.n_acc[m] >= 2u, m);
    }
}

// Round r, first kernel: the best track of every shared identifier.
__global__ __launch_bounds__(kAmbBlock) void k_amb_best(amb_params P, uint32_t r) {
    amb_ctl* c = P.ctl;
    if (c->done != 0u) return;
    if (r > 0u && c->removed[(r - 1u) & 1u] == 0u) {
        if (blockIdx.x == 0u && threadIdx.x == 0u) {
            c->done = 1u;
            c->last_round = r;
        }
        return;
    }
    if (blockIdx.x == 0u && threadIdx.x == 0u) {
        c->n_cand[(r + 1u) & 1u] = 0u;  // last read as round r - 1's list
        c->removed[r & 1u] = 0u;
    }
    const uint32_t nh = c->n_hot[r & 1u];
    const uint32_t* H = P.hot[r & 1u];
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < nh; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        bool keep = false;
        uint32_t m = 0u;
        if (i < nh) {
            m = H[i];
            if (P.n_acc[m] >= 2u) {
                keep = true;
                track_key best{0.f, 0.f, kNone};
                bool have = false;
                for (uint32_t q = P.list_off[m]; q < P.list_off[m + 1u]; ++q) {
                    const uint32_t u = P.list[q];
                    if (P.status[u] != kAlive) continue;
                    const track_key k = key_of(P, u);
                    if (!have || worse(best, k)) {
                        best = k;
                        have = true;
                    }
                }
                if (have) {
                    P.best_of[m] = best.t;
                    P.protect[best.t] = r + 1u;
                }
            }
        }
        wave_append(&c->n_hot[(r + 1u) & 1u], P.hot[(r + 1u) & 1u], keep, m);
    }
}

// Round r, second kernel: remove every candidate that is best on none of its
// identifiers.  Only protected tracks have n_shared lowered here, and a
// protected track is never removed, so the removal test reads stable values.
__global__ __launch_bounds__(kAmbBlock) void k_amb_remove(amb_params P, uint32_t r) {
    amb_ctl* c = P.ctl;
    if (c->done != 0u) return;
    if (blockIdx.x == 0u && threadIdx.x == 0u) {
        c->n_hot[r & 1u] = 0u;  // read as this round's list by k_amb_best only
    }
    const uint32_t nt = c->n_cand[r & 1u];
    const uint32_t* T = P.cand[r & 1u];
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < nt; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        bool keep = false, removing = false;
        uint32_t t = 0u;
        if (i < nt) {
            t = T[i];
            if (P.status[t] == kAlive && P.n_shared[t] >= 1u) {
                if (P.protect[t] == r + 1u) {
                    keep = true;
                } else {
                    removing = true;
                }
            }
        }
        wave_append(&c->n_cand[(r + 1u) & 1u], P.cand[(r + 1u) & 1u], keep, t);
        wave_count(&c->removed[r & 1u], removing);
        wave_count(&c->total_removed, removing);
        if (removing) {
            P.status[t] = kDead;
            const uint32_t b = P.off[t], e = P.off[t + 1];
            for (uint32_t j = b; j < e; ++j) {
                if (!first_occurrence(P.ids, b, j)) continue;
                const uint32_t id = P.ids[j] - P.id_base;
                if (atomicSub(&P.n_acc[id], 1u) == 2u) {
                    // the identifier's one remaining track is its best
                    const uint32_t s = P.best_of[id];
                    if (s < P.n_tracks) atomicSub(&P.n_shared[s], multiplicity(P, s, P.ids[j]));
                }
            }
        }
    }
}

// One workgroup, one removal per iteration, as the reference iterates:
// the worst candidate and the largest n_shared, the stop test, the removal.
__global__ __launch_bounds__(kSeqBlock) void k_amb_sequential(amb_params P, uint32_t max_shared,
                                                              uint32_t max_iter) {
    __shared__ uint32_t s_t[kSeqBlock];
    __shared__ uint32_t s_max[kSeqBlock];
    const uint32_t tid = threadIdx.x;
    const uint32_t nt = P.ctl->n_cand[0];
    const uint32_t* T = P.cand[0];
    uint32_t removed = 0u;
    for (uint32_t it = 0; it < max_iter; ++it) {
        uint32_t wt = kNone, mx = 0u;
        track_key wk{0.f, 0.f, kNone};
        for (uint32_t i = tid; i < nt; i += kSeqBlock) {
            const uint32_t t = T[i];
            if (P.status[t] != kAlive) continue;
            const uint32_t ns = P.n_shared[t];
            if (ns == 0u) continue;
            mx = ns > mx ? ns : mx;
            const track_key k = key_of(P, t);
            if (wt == kNone || worse(k, wk)) {
                wk = k;
                wt = t;
            }
        }
        s_t[tid] = wt;
        s_max[tid] = mx;
        __syncthreads();
        for (uint32_t s = kSeqBlock / 2u; s > 0u; s >>= 1u) {
            if (tid < s) {
                const uint32_t a = s_t[tid], b = s_t[tid + s];
                if (b != kNone && (a == kNone || worse(key_of(P, b), key_of(P, a)))) {
                    s_t[tid] = b;
                }
                s_max[tid] = s_max[tid + s] > s_max[tid] ? s_max[tid + s] : s_max[tid];
            }
            __syncthreads();
        }
        const uint32_t W = s_t[0];
        const uint32_t M = s_max[0];
        __syncthreads();
        if (W == kNone || M < max_shared) break;
        const uint32_t b = P.off[W], e = P.off[W + 1];
        for (uint32_t j = b + tid; j < e; j += kSeqBlock) {
            if (!first_occurrence(P.ids, b, j)) continue;
            const uint32_t id = P.ids[j] - P.id_base;
            if (atomicSub(&P.n_acc[id], 1u) == 2u) {
                uint32_t s = kNone;
                for (uint32_t q = P.list_off[id]; q < P.list_off[id + 1u]; ++q) {
                    const uint32_t u = P.list[q];
                    if (u != W && P.status[u] == kAlive) {
                        s = u;
                        break;
                    }
                }
                if (s != kNone) atomicSub(&P.n_shared[s], multiplicity(P, s, P.ids[j]));
            }
        }
        __syncthreads();
        if (tid == 0u) P.status[W] = kDead;
        ++removed;
        __syncthreads();
    }
    if (tid == 0u) P.ctl->total_removed = removed;
}

__global__ __launch_bounds__(kAmbBlock) void k_amb_output(amb_params P, uint8_t* accepted) {
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < P.n_tracks; b0 += gridDim.x * blockDim.x) {
        const uint32_t t = b0 + threadIdx.x;
        const bool a = t < P.n_tracks && P.status[t] == kAlive;
        if (t < P.n_tracks) accepted[t] = a ? 1u : 0u;
        wave_count(&P.ctl->n_accepted, a);
    }
}

// Per event of a fitted batch: identifiers = event measurement offset +
// the constituent's event-relative index.
__global__ __launch_bounds__(kAmbBlock) void k_amb_fitted_ids(const uint32_t* off, const uint32_t* cons,
                                                              const uint32_t* ev_track_off,
                                                              const uint32_t* ev_meas_off,
                                                              uint32_t n_events, uint32_t n_tracks,
                                                              uint32_t* ids) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_tracks; t += gridDim.x * blockDim.x) {
        uint32_t lo = 0u, hi = n_events;  // largest e with ev_track_off[e] <= t
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) / 2u;
            if (ev_track_off[mid] <= t) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        const uint32_t mo = ev_meas_off[lo];
        for (uint32_t j = off[t]; j < off[t + 1]; ++j) ids[j] = mo + cons[j];
    }
}

uint32_t grid_for(uint64_t items) {
    const uint64_t b = (items + kAmbBlock - 1u) / kAmbBlock;
    return static_cast<uint32_t>(std::max<uint64_t>(1u, std::min<uint64_t>(b, 8192u)));
}

}  // namespace
}  // namespace tamd

using tamd::ckf_error;
using tamd::guarded;
using tamd::hip_check;

struct ckf_resolver {
    ckf_ambiguity_config cfg{};
    hipStream_t stream = nullptr;
    ckf_resolver_stats stats{};
    tamd::dbuf status, n_shared, protect, cand0, cand1, n_acc, list_off, fill, list, best_of, hot0,
        hot1, sums, ctl, in_off, in_ids, in_pval, out_acc, fit_ids, fit_ev_off;
    tamd::amb_ctl* h_ctl = nullptr;  // pinned
    ~ckf_resolver() {
        tamd::dbuf* bs[] = {&status, &n_shared, &protect, &cand0,  &cand1,   &n_acc,   &list_off,
                            &fill,   &list,     &best_of, &hot0,   &hot1,    &sums,    &ctl,
                            &in_off, &in_ids,   &in_pval, &out_acc, &fit_ids, &fit_ev_off};
        for (tamd::dbuf* b : bs) b->release();
        if (h_ctl != nullptr) (void)hipHostFree(h_ctl);
    }
};

namespace {

void read_ctl(ckf_resolver& R) {
    hip_check(hipMemcpyAsync(R.h_ctl, R.ctl.p, sizeof(tamd::amb_ctl), hipMemcpyDeviceToHost, R.stream),
              "control readback");
    hip_check(hipStreamSynchronize(R.stream), "resolver");
}

// Build the identifiers' track lists and the first candidate list.
void setup(ckf_resolver& R, tamd::amb_params& P, uint32_t n_entries) {
    using namespace tamd;
    const hipStream_t s = R.stream;
    const uint32_t B = std::max(P.id_bound, 1u);
    hip_check(hipMemsetAsync(R.ctl.p, 0, sizeof(amb_ctl), s), "memset");
    hip_check(hipMemsetAsync(P.n_acc, 0, 4u * size_t(B), s), "memset");
    hip_check(hipMemsetAsync(P.fill, 0, 4u * size_t(B), s), "memset");
    hip_check(hipMemsetAsync(P.best_of, 0xff, 4u * size_t(B), s), "memset");
    hipLaunchKernelGGL(k_amb_count, dim3(grid_for(P.n_tracks)), dim3(kAmbBlock), 0, s, P);
    hip_check(hipGetLastError(), "count");
    exclusive_scan(P.n_acc, P.list_off, R.sums.as<uint32_t>(), B, &P.ctl->scan_total, s);
    read_ctl(R);
    if (R.h_ctl->bad_input & 1u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "constituent offsets decrease or pass offsets[n_tracks]");
    }
    if (R.h_ctl->bad_input & 2u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "a measurement identifier is not below id_bound");
    }
    if (R.h_ctl->scan_total > n_entries) {
        throw ckf_error(CKF_ERR_INTERNAL, "identifier lists exceed the constituents");
    }
    hipLaunchKernelGGL(k_amb_fill, dim3(grid_for(P.n_tracks)), dim3(kAmbBlock), 0, s, P);
    hipLaunchKernelGGL(k_amb_hot0, dim3(grid_for(B)), dim3(kAmbBlock), 0, s, P);
    hip_check(hipGetLastError(), "fill");
}

void run_rounds(ckf_resolver& R, tamd::amb_params& P) {
    using namespace tamd;
    const hipStream_t s = R.stream;
    // a round's grid covers its lists (they only shrink); kernels stride over the rest
    const uint32_t gb = grid_for(std::max(P.id_bound, 1u));
    const uint32_t gr = grid_for(std::max(P.n_tracks, 1u));
    constexpr uint32_t kChunk = 4;
    for (uint32_t r = 0;; ++r) {
        hipLaunchKernelGGL(k_amb_best, dim3(gb), dim3(kAmbBlock), 0, s, P, r);
        hipLaunchKernelGGL(k_amb_remove, dim3(gr), dim3(kAmbBlock), 0, s, P, r);
        hip_check(hipGetLastError(), "round");
        if (r % kChunk == kChunk - 1u) {
            read_ctl(R);
            if (R.h_ctl->done) break;
            if (r > P.n_tracks + 2u * kChunk) {
                throw ckf_error(CKF_ERR_INTERNAL, "ambiguity rounds did not converge");
            }
        }
    }
}

// The operator on device arrays; fills R.stats and d_accepted.
void resolve_device(ckf_resolver& R, uint32_t n, const uint32_t* d_off, const uint32_t* d_ids,
                    const float* d_pval, uint32_t id_base, uint32_t id_bound, uint32_t n_entries,
                    uint8_t* d_accepted) {
    using namespace tamd;
    const hipStream_t s = R.stream;
    const uint32_t N = std::max(n, 1u), B = std::max(id_bound, 1u);
    R.status.ensure(4u * size_t(N));
    R.n_shared.ensure(4u * size_t(N));
    R.protect.ensure(4u * size_t(N));
    R.cand0.ensure(4u * size_t(N));
    R.cand1.ensure(4u * size_t(N));
    R.n_acc.ensure(4u * size_t(B));
    R.list_off.ensure(4u * (size_t(B) + 1u));
    R.fill.ensure(4u * size_t(B));
    R.list.ensure(4u * std::max<size_t>(n_entries, 1u));
    R.best_of.ensure(4u * size_t(B));
    R.hot0.ensure(4u * size_t(B));
    R.hot1.ensure(4u * size_t(B));
    R.sums.ensure(4u * size_t(scan_sums_size(B)));
    R.ctl.ensure(sizeof(amb_ctl));
    if (R.h_ctl == nullptr) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&R.h_ctl), sizeof(amb_ctl)), "hipHostMalloc");
    }
    amb_params P{};
    P.n_tracks = n;
    P.id_base = id_base;
    P.id_bound = id_bound;
    P.min_meas = R.cfg.min_meas_per_track;
    P.n_entries = n_entries;
    P.off = d_off;
    P.ids = d_ids;
    P.pval = d_pval;
    P.status = R.status.as<uint32_t>();
    P.n_shared = R.n_shared.as<uint32_t>();
    P.protect = R.protect.as<uint32_t>();
    P.cand[0] = R.cand0.as<uint32_t>();
    P.cand[1] = R.cand1.as<uint32_t>();
    P.n_acc = R.n_acc.as<uint32_t>();
    P.list_off = R.list_off.as<uint32_t>();
    P.fill = R.fill.as<uint32_t>();
    P.list = R.list.as<uint32_t>();
    P.best_of = R.best_of.as<uint32_t>();
    P.hot[0] = R.hot0.as<uint32_t>();
    P.hot[1] = R.hot1.as<uint32_t>();
    P.ctl = R.ctl.as<amb_ctl>();

    ckf_resolver_stats st{};
    st.n_tracks = n;
    setup(R, P, n_entries);
    bool sequential = R.cfg.max_shared_meas != 1u;
    if (!sequential) {
        run_rounds(R, P);
        read_ctl(R);
        st.n_rounds = R.h_ctl->last_round > 0u ? R.h_ctl->last_round - 1u : 0u;
        if (R.h_ctl->total_removed > R.cfg.max_iterations) {
            // the reference stops after max_iterations removals: redo one at a time
            sequential = true;
            st.n_rounds = 0u;
            setup(R, P, n_entries);
        }
    }
    if (sequential) {
        hipLaunchKernelGGL(k_amb_sequential, dim3(1), dim3(kSeqBlock), 0, s, P, R.cfg.max_shared_meas,
                           R.cfg.max_iterations);
        hip_check(hipGetLastError(), "sequential");
    }
    hipLaunchKernelGGL(k_amb_output, dim3(grid_for(N)), dim3(kAmbBlock), 0, s, P, d_accepted);
    hip_check(hipGetLastError(), "output");
    read_ctl(R);
    st.n_candidates = R.h_ctl->n_candidates;
    st.n_removed = R.h_ctl->total_removed;
    st.n_accepted = R.h_ctl->n_accepted;
    st.sequential = sequential ? 1u : 0u;
    R.stats = st;
}

void check_host_offsets(const ckf_resolver_tracks& T) {
    if (T.constituent_offsets[0] != 0u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "constituent offsets must start at 0");
    }
    for (uint32_t t = 0; t < T.n_tracks; ++t) {
        if (T.constituent_offsets[t + 1] < T.constituent_offsets[t]) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "constituent offsets decrease");
        }
    }
}

void check_tracks_struct(const ckf_resolver_tracks* T) {
    if (T == nullptr || (T->n_tracks > 0u && (T->constituent_offsets == nullptr || T->pval == nullptr))) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null track arrays");
    }
}

}  // namespace

extern "C" {

void ckf_ambiguity_config_default(ckf_ambiguity_config* cfg) {
    if (cfg == nullptr) return;
    cfg->min_meas_per_track = 3u;
    cfg->max_iterations = std::numeric_limits<uint32_t>::max();
    cfg->max_shared_meas = 1u;
}

ckf_status ckf_resolver_create(const ckf_ambiguity_config* cfg, ckf_resolver** out) {
    return guarded([&] {
        if (cfg == nullptr || out == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        if (cfg->max_shared_meas == 0u) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "max_shared_meas must be larger than zero");
        }
        auto* r = new ckf_resolver();
        r->cfg = *cfg;
        *out = r;
    });
}

ckf_status ckf_resolver_destroy(ckf_resolver* r) {
    return guarded([&] { delete r; });
}

ckf_status ckf_resolver_set_stream(ckf_resolver* r, void* hip_stream) {
    return guarded([&] {
        if (r == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null resolver");
        r->stream = static_cast<hipStream_t>(hip_stream);
    });
}

ckf_status ckf_resolver_last_stats(const ckf_resolver* r, ckf_resolver_stats* stats) {
    return guarded([&] {
        if (r == nullptr || stats == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        *stats = r->stats;
    });
}

ckf_status ckf_resolve_ambiguities(ckf_resolver* r, const ckf_resolver_tracks* tracks,
                                   uint8_t* accepted, uint32_t* n_accepted) {
    return guarded([&] {
        if (r == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null resolver");
        check_tracks_struct(tracks);
        const uint32_t n = tracks->n_tracks;
        if (n > 0u && accepted == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null output");
        if (n == 0u) {
            r->stats = ckf_resolver_stats{};
            if (n_accepted) *n_accepted = 0u;
            return;
        }
        check_host_offsets(*tracks);
        const uint32_t ne = tracks->constituent_offsets[n];
        if (ne > 0u && tracks->measurement_id == nullptr) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null measurement identifiers");
        }
        const hipStream_t s = r->stream;
        r->in_off.ensure(4u * (size_t(n) + 1u));
        r->in_ids.ensure(4u * std::max<size_t>(ne, 1u));
        r->in_pval.ensure(4u * size_t(n));
        r->out_acc.ensure(size_t(n));
        hip_check(hipMemcpyAsync(r->in_off.p, tracks->constituent_offsets, 4u * (size_t(n) + 1u),
                                 hipMemcpyHostToDevice, s), "upload offsets");
        if (ne > 0u) {
            hip_check(hipMemcpyAsync(r->in_ids.p, tracks->measurement_id, 4u * size_t(ne),
                                     hipMemcpyHostToDevice, s), "upload identifiers");
        }
        hip_check(hipMemcpyAsync(r->in_pval.p, tracks->pval, 4u * size_t(n), hipMemcpyHostToDevice, s),
                  "upload p-values");
        resolve_device(*r, n, r->in_off.as<uint32_t>(), r->in_ids.as<uint32_t>(),
                       r->in_pval.as<float>(), 0u, tracks->id_bound, ne, r->out_acc.as<uint8_t>());
        hip_check(hipMemcpyAsync(accepted, r->out_acc.p, n, hipMemcpyDeviceToHost, s), "download");
        hip_check(hipStreamSynchronize(s), "resolver");
        if (n_accepted) *n_accepted = r->stats.n_accepted;
    });
}

ckf_status ckf_resolve_ambiguities_device(ckf_resolver* r, const ckf_resolver_tracks* tracks,
                                          uint8_t* d_accepted, uint32_t* n_accepted) {
    return guarded([&] {
        if (r == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null resolver");
        check_tracks_struct(tracks);
        const uint32_t n = tracks->n_tracks;
        if (n > 0u && d_accepted == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null output");
        if (n == 0u) {
            r->stats = ckf_resolver_stats{};
            if (n_accepted) *n_accepted = 0u;
            return;
        }
        uint32_t ends[2] = {0u, 0u};
        hip_check(hipMemcpyAsync(&ends[0], tracks->constituent_offsets, 4u, hipMemcpyDeviceToHost,
                                 r->stream), "read offsets");
        hip_check(hipMemcpyAsync(&ends[1], tracks->constituent_offsets + n, 4u, hipMemcpyDeviceToHost,
                                 r->stream), "read offsets");
        hip_check(hipStreamSynchronize(r->stream), "resolver");
        if (ends[0] != 0u) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "constituent offsets must start at 0");
        if (ends[1] > 0u && tracks->measurement_id == nullptr) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null measurement identifiers");
        }
        // the count kernel flags an offset that decreases or passes
        // offsets[n] before any kernel reads the identifiers
        resolve_device(*r, n, tracks->constituent_offsets, tracks->measurement_id, tracks->pval, 0u,
                       tracks->id_bound, ends[1], d_accepted);
        if (n_accepted) *n_accepted = r->stats.n_accepted;
    });
}

ckf_status ckf_resolve_fitted(ckf_resolver* r, const ckf_device_event* ev,
                              const ckf_device_fitted* fitted, uint8_t* accepted, uint32_t* n_accepted) {
    return guarded([&] {
        if (r == nullptr || ev == nullptr || fitted == nullptr) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        const uint32_t n = fitted->n_tracks, ns = fitted->n_states;
        if (fitted->ev_track_off.size() != size_t(ev->n_events) + 1u) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "fitted tracks and events do not match");
        }
        if (n > 0u && accepted == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null output");
        if (n == 0u) {
            r->stats = ckf_resolver_stats{};
            if (n_accepted) *n_accepted = 0u;
            return;
        }
        const hipStream_t s = r->stream;
        r->fit_ids.ensure(4u * std::max<size_t>(ns, 1u));
        r->fit_ev_off.ensure(4u * fitted->ev_track_off.size());
        r->out_acc.ensure(size_t(n));
        hip_check(hipMemcpyAsync(r->fit_ev_off.p, fitted->ev_track_off.data(),
                                 4u * fitted->ev_track_off.size(), hipMemcpyHostToDevice, s),
                  "event offsets");
        const uint32_t* d_off = fitted->offsets.as<uint32_t>();
        hipLaunchKernelGGL(tamd::k_amb_fitted_ids, dim3(tamd::grid_for(n)), dim3(tamd::kAmbBlock), 0, s,
                           d_off, fitted->cons.as<uint32_t>(), r->fit_ev_off.as<uint32_t>(),
                           ev->d_meas_off, ev->n_events, n, r->fit_ids.as<uint32_t>());
        hip_check(hipGetLastError(), "identifiers");
        const uint32_t* ids = r->fit_ids.as<uint32_t>();
        const float* pv = fitted->pval.as<float>();
        uint8_t* out = r->out_acc.as<uint8_t>();
        const bool batch_exact = r->cfg.max_shared_meas == 1u &&
                                 r->cfg.max_iterations == std::numeric_limits<uint32_t>::max();
        if (batch_exact || ev->n_events == 1u) {
            // events share no measurement: one run over the batch equals a
            // run per event when the rounds decide
            resolve_device(*r, n, d_off, ids, pv, 0u, ev->n_meas, ns, out);
        } else {
            // the stop test and the iteration limit apply per event
            ckf_resolver_stats sum{};
            for (uint32_t e = 0; e < ev->n_events; ++e) {
                const uint32_t t0 = fitted->ev_track_off[e], t1 = fitted->ev_track_off[e + 1];
                if (t1 == t0) continue;
                uint32_t ends[2];
                hip_check(hipMemcpyAsync(&ends[0], d_off + t0, 4u, hipMemcpyDeviceToHost, s), "offsets");
                hip_check(hipMemcpyAsync(&ends[1], d_off + t1, 4u, hipMemcpyDeviceToHost, s), "offsets");
                hip_check(hipStreamSynchronize(s), "resolver");
                resolve_device(*r, t1 - t0, d_off + t0, ids, pv + t0, ev->meas_off[e],
                               ev->meas_off[e + 1] - ev->meas_off[e], ends[1], out + t0);
                sum.n_candidates += r->stats.n_candidates;
                sum.n_removed += r->stats.n_removed;
                sum.n_accepted += r->stats.n_accepted;
                sum.n_rounds = std::max(sum.n_rounds, r->stats.n_rounds);
                sum.sequential |= r->stats.sequential;
            }
            sum.n_tracks = n;
            r->stats = sum;
        }
        hip_check(hipMemcpyAsync(accepted, out, n, hipMemcpyDeviceToHost, s), "download");
        hip_check(hipStreamSynchronize(s), "resolver");
        if (n_accepted) *n_accepted = r->stats.n_accepted;
    });
}

}  // extern "C"
