// CKF step kernels for CDNA (gfx950).
//
// Per step (counts live in the device step control block):
//   k_find        one lane per input branch: chi2 over the surface's
//                 measurements, best-N selection by (chi2, measurement index),
//                 gain-matrix update, hole links; links are compacted in input
//                 order with a wave scan + decoupled look-back
//                 (traccc host: details/combinatorial_kalman_filter.hpp:174-364)
//   k_dedup_*     duplicate removal grouped by last measurement
//                 (traccc host :371-488)
//   k_propagate   one lane per link: per-seed cap, hole limits, RK4 transport
//                 to the next sensitive surface, tips; next-step inputs and
//                 tips compacted with ballots + look-back (traccc host :493-627)
//   k_advance     one lane: advances the control block
// build_tracks after the loop (traccc host :637-715).
//
// Work distribution: every kernel is a persistent grid whose waves take
// 64-item partitions from an atomic ticket (device work queue), so uneven
// per-branch cost (measurements per surface, RK steps per crossing) balances
// across the chip.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "device_types.hpp"
#include "kalman_device.hpp"
#include "features.hpp"
#include "kernels.hpp"
#include "../scorer/mlp_device.hpp"
#include "state_io.hpp"
#include "wave.hpp"

namespace tamd {

namespace {

constexpr int kBlock = 256;

__device__ inline void set_overflow(step_ctl* ctl, uint32_t bit) { atomicOr(&ctl->overflow, bit); }

__device__ inline dlink load_link(const ckf_params& P, uint32_t g) { return P.links[g]; }

}  // namespace

// ---------------------------------------------------------------------------
// find_tracks
// ---------------------------------------------------------------------------
// A team of kT lanes handles one input branch: the lanes split the
// surface's measurements, reduce the best (chi2, index) candidate with
// shuffles, and run the gain update redundantly (so the team stays uniform);
// the team leader writes.  A wave holds kWave / kT inputs.  Each lane reads
// kBatch measurements per round with all loads issued before the chi2
// arithmetic, so a round costs one memory latency.  The plain find runs
// two-lane teams (its gain updates are redundant over the team, and the loc0
// window leaves few measurements to split); the scorer list, which updates
// nothing, runs four-lane teams.
static_assert(sizeof(line_u32) == kXcdCounterStride * 4u &&
                  sizeof(step_ctl::xcd_ticket[0]) == kXcds * sizeof(line_u32),
              "per-XCD counters: one per XCD, 128 B apart");
constexpr int kFindTeam = 2;
constexpr int kListTeam = 4;
// scorer lists: up to kListTeam * kListSlots candidates per input
constexpr int kListSlots = 4;
constexpr int kBatch = 4;

// One pass of a team over the surface's measurements [lo, up): the best
// (chi2, index) candidate after (last_c, last_m) (every candidate when
// !have_last) that passes the cut and the step-0 rule.  kInvalid when none.
// With count_pass, n_pass receives the number of measurements passing the cut
// (on every lane of the team).
template <int kT>
__device__ inline void scan_next(const ckf_params& P, const float* x, const float* c, uint32_t lo,
                                 uint32_t up, bool rule0, bool have_last, float last_c,
                                 uint32_t last_m, uint32_t tl, bool count_pass, float& best_c,
                                 uint32_t& best_m, uint32_t& n_pass) {
    best_c = 0.f;
    best_m = kInvalid;
    uint32_t np = 0u;
    for (uint32_t m0 = lo + tl; m0 < up; m0 += kT * kBatch) {
        meas_data mb[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const uint32_t m = m0 + static_cast<uint32_t>(u * kT);
            mb[u] = load_meas(P.meas_pos, P.meas_var, P.meas_info, m < up ? m : lo);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const uint32_t m = m0 + static_cast<uint32_t>(u * kT);
            const float chi2 = predicted_chi2(mb[u], x, c);
            // outside the cut (a NaN never yields a link either)
            bool take = m < up && chi2 <= P.chi2_max && chi2 >= 0.f;
            take = take && !(rule0 && chi2 != 0.f);
            np += take ? 1u : 0u;
            take = take && !(have_last && (chi2 < last_c || (chi2 == last_c && m <= last_m)));
            if (take && (best_m == kInvalid || chi2 < best_c || (chi2 == best_c && m < best_m))) {
                best_c = chi2;
                best_m = m;
            }
        }
    }
    // team minimum of (chi2, index); kInvalid sorts last
#pragma unroll
    for (int d = kT / 2; d >= 1; d >>= 1) {
        const float oc = __shfl_xor(best_c, d, kT);
        const uint32_t om = __shfl_xor(best_m, d, kT);
        const bool take = (om != kInvalid) &&
                          (best_m == kInvalid || oc < best_c || (oc == best_c && om < best_m));
        if (take) {
            best_c = oc;
            best_m = om;
        }
    }
    if (count_pass) {
#pragma unroll
        for (int d = kT / 2; d >= 1; d >>= 1) np += __shfl_xor(np, d, kT);
        n_pass = np;
    }
}

// The input branch a team works on at this step: its predicted state, its
// previous link's bookkeeping and the surface's measurement range.
struct find_input {
    float x[6], c[21];
    uint32_t sf, seed, prev_g, skip, consec, ndf_prev, in_slot, lo, up;
    float chi2_prev;
    float win_var;  // the event's largest loc0 variance, < 0: no window search
    bool rule0;
};

// A superset [lo, up) of the surface's measurements whose loc0 can lie in
// [a, b], for measurements sorted by loc0: cooperative 16-ary narrowing by
// the team (kT lanes x kBatch probes per round, one memory latency each)
// of the first index >= a and of the first index > b.
template <int kT>
__device__ inline void window_range(const ckf_params& P, float a, float b, uint32_t tl,
                                    uint32_t& lo, uint32_t& up) {
    constexpr uint32_t kProbes = kT * kBatch;
    for (int bound = 0; bound < 2; ++bound) {
        const float key = bound == 0 ? a : b;
        uint32_t L = lo, H = up;  // the wanted index lies in [L, H]
        while (H - L > kProbes) {
            const uint32_t span = H - L;
            uint32_t below = 0u;
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t k = static_cast<uint32_t>(u * kT) + tl;
                const uint32_t pos = L + ((k + 1u) * span) / (kProbes + 1u);
                const float v = P.meas_pos[pos].x;
                below += (bound == 0 ? (v < key) : !(v > key)) ? 1u : 0u;
            }
#pragma unroll
            for (int d = kT / 2; d >= 1; d >>= 1) below += __shfl_xor(below, d, kT);
            // probes are sorted: `below` of them precede the wanted index
            const uint32_t nL = below == 0u ? L : L + (below * span) / (kProbes + 1u) + 1u;
            const uint32_t nH = below == kProbes ? H : L + ((below + 1u) * span) / (kProbes + 1u);
            L = nL;
            H = nH;
        }
        if (bound == 0) {
            lo = L;
        } else {
            up = H;
        }
    }
}

// Candidates of one input in (chi2, index) order from a single pass over the
// surface's measurements (its loc0 window when the event allows one): every
// lane keeps the kLaneKeep best passing candidates of its share in registers
// (sorted), and next() merges the team's lists with one team minimum per
// candidate.  When some lane of the team saw more passing candidates than it
// keeps, the stream falls back to one scan_next pass over the same range per
// candidate (the order is the same either way).
template <int kLaneKeep, int kT>
struct cand_stream {
    static constexpr int kKeep = kLaneKeep;
    static constexpr int kTeamSize = kT;
    float lc[kLaneKeep];
    uint32_t lm[kLaneKeep];
    uint32_t lo, up;    // the measurements searched (the loc0 window when taken)
    uint32_t head;      // next own list entry
    uint32_t n_pass;    // team: measurements passing the cut
    bool fallback;      // team-uniform
    bool have_last;
    float last_c;
    uint32_t last_m;
};

template <typename Stream>
__device__ __forceinline__ void stream_init(const ckf_params& P, const find_input& in, uint32_t tl,
                                   Stream& S) {
    constexpr int kLaneKeep = Stream::kKeep;
    constexpr int kT = Stream::kTeamSize;
    const float inf = __int_as_float(0x7f800000);
#pragma unroll
    for (int j = 0; j < kLaneKeep; ++j) {
        S.lc[j] = inf;
        S.lm[j] = kInvalid;
    }
    uint32_t np = 0u;
    // a measurement passing the cut has r0^2 <= chi2_max (P00 + V00)
    // (chi2 >= r0^2 / R00 for any 2x2 SPD R): outside that loc0 window none
    // can pass, so on a dense surface sorted by loc0 only the window is read
    uint32_t lo = in.lo, up = in.up;
    if (in.win_var >= 0.f && up - lo > static_cast<uint32_t>(kT * kBatch)) {
        const float w = sqrtf(P.chi2_max * (in.c[sym_idx(0, 0)] + in.win_var)) * 1.01f + 1e-3f;
        if (w == w) {
            window_range<kT>(P, in.x[e_loc0] - w, in.x[e_loc0] + w, tl, lo, up);
        }
    }
    for (uint32_t m0 = lo + tl; m0 < up; m0 += kT * kBatch) {
        meas_data mb[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const uint32_t m = m0 + static_cast<uint32_t>(u * kT);
            mb[u] = load_meas(P.meas_pos, P.meas_var, P.meas_info, m < up ? m : lo);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const uint32_t m = m0 + static_cast<uint32_t>(u * kT);
            const float chi2 = predicted_chi2(mb[u], in.x, in.c);
            bool take = m < up && chi2 <= P.chi2_max && chi2 >= 0.f;
            take = take && !(in.rule0 && chi2 != 0.f);
            if (take) {
                ++np;
                // insertion into the list sorted by (chi2, index)
                float cc = chi2;
                uint32_t cm = m;
#pragma unroll
                for (int j = 0; j < kLaneKeep; ++j) {
                    if (cc < S.lc[j] || (cc == S.lc[j] && cm < S.lm[j])) {
                        const float tc = S.lc[j];
                        const uint32_t tm = S.lm[j];
                        S.lc[j] = cc;
                        S.lm[j] = cm;
                        cc = tc;
                        cm = tm;
                    }
                }
            }
        }
    }
    uint32_t over = np > static_cast<uint32_t>(kLaneKeep) ? 1u : 0u;
#pragma unroll
    for (int d = kT / 2; d >= 1; d >>= 1) {
        np += __shfl_xor(np, d, kT);
        over |= __shfl_xor(over, d, kT);
    }
    S.n_pass = np;
    S.lo = lo;
    S.up = up;
    S.fallback = over != 0u;
    S.head = 0u;
    S.have_last = false;
    S.last_c = 0.f;
    S.last_m = 0u;
}

template <typename Stream>
__device__ __forceinline__ void stream_next(const ckf_params& P, const find_input& in, uint32_t tl,
                                   Stream& S, float& best_c, uint32_t& best_m) {
    constexpr int kLaneKeep = Stream::kKeep;
    constexpr int kT = Stream::kTeamSize;
    if (S.fallback) {
        uint32_t unused = 0u;
        scan_next<kT>(P, in.x, in.c, S.lo, S.up, in.rule0, S.have_last, S.last_c, S.last_m, tl,
                  false, best_c, best_m, unused);
    } else {
        // own head: a select chain keeps the list in registers
        float c = S.lc[kLaneKeep - 1];
        uint32_t m = S.lm[kLaneKeep - 1];
#pragma unroll
        for (int j = kLaneKeep - 2; j >= 0; --j) {
            if (S.head == static_cast<uint32_t>(j)) {
                c = S.lc[j];
                m = S.lm[j];
            }
        }
        if (S.head >= static_cast<uint32_t>(kLaneKeep)) {
            m = kInvalid;
        }
        best_c = c;
        best_m = m;
        // team minimum of (chi2, index); kInvalid sorts last
#pragma unroll
        for (int d = kT / 2; d >= 1; d >>= 1) {
            const float oc = __shfl_xor(best_c, d, kT);
            const uint32_t om = __shfl_xor(best_m, d, kT);
            const bool take = (om != kInvalid) &&
                              (best_m == kInvalid || oc < best_c || (oc == best_c && om < best_m));
            if (take) {
                best_c = oc;
                best_m = om;
            }
        }
        if (best_m != kInvalid && best_m == m) {
            ++S.head;  // the owner moves to its next entry
        }
    }
    if (best_m != kInvalid) {
        S.have_last = true;
        S.last_c = best_c;
        S.last_m = best_m;
    }
}

// kRanges = false (the scored update, which scans nothing) skips the
// event's measurement range and window lookups.
template <bool kRanges = true>
__device__ inline void load_input(const ckf_params& P, uint32_t step, uint32_t i,
                                  uint32_t prev_base, find_input& in) {
    const uint32_t in_par = (step + 1u) & 1u;
    in.prev_g = kInvalid;
    in.skip = in.consec = in.ndf_prev = in.in_slot = 0u;
    in.chi2_prev = 0.f;
    if (step == 0u) {
        in.seed = i;
        load_seed(P, i, in.x, in.c, in.sf);
    } else {
        in.in_slot = P.live[i];
        in.prev_g = prev_base + in.in_slot;
        const dlink L = load_link(P, in.prev_g);
        in.seed = L.seed;
        in.skip = link_skipped(L);
        in.consec = link_consec(L);
        in.chi2_prev = L.chi2_sum;
        in.ndf_prev = L.ndf_sum;
        load_soa(P.state[in_par], P.state_cap, in.in_slot, in.x, in.c, in.sf);
    }
    in.lo = in.up = 0u;
    in.win_var = -1.f;
    if constexpr (kRanges) {
        const uint32_t ev = P.seed_event[in.seed];
        const uint32_t* rg = P.meas_ranges + static_cast<size_t>(ev) * (P.n_surfaces + 1u);
        in.lo = rg[in.sf];
        in.up = rg[in.sf + 1u];
        in.win_var = P.event_win_var[ev];
    }
    in.rule0 = (step == 0u) && !(P.sfs[in.sf].thickness_x0 > 0.f);
}

// Link bookkeeping handed from k_find to k_find_emit (see stage_head).
__device__ inline stage_head make_stage_head(const find_input& in, uint32_t step, uint32_t i) {
    stage_head h;
    h.prev = in.prev_g;
    h.seed = in.seed;
    h.hist = in.skip | (in.consec << 8);
    h.hole_ref = (step == 0u) ? ((kRefSeed << kRefKindShift) | i)
                              : ((kRefInput << kRefKindShift) | in.in_slot);
    h.chi2_sum = in.chi2_prev;
    h.ndf_sum = in.ndf_prev;
    h.pad[0] = h.pad[1] = 0u;
    return h;
}

// Measurement indices stay below 2^31 (the batch's measurement count is a
// 32-bit size and far smaller), so the top bit carries the dimension.
__device__ inline stage_cand make_stage_cand(uint32_t m, float chi2, uint32_t dims) {
    stage_cand c;
    c.meas = m | (dims == 2u ? 0x80000000u : 0u);
    c.chi2 = chi2;
    return c;
}

// Filtered-state slot of an input's k-th kept candidate: plane k of in_cap
// slots, so the common case of one kept candidate per input fills
// consecutive slots (coalesced SoA stores here and loads in k_propagate).
// Past flt_cap it returns kInvalid, which every caller treats as overflow.
__device__ inline uint32_t flt_slot(const ckf_params& P, uint32_t i, uint32_t k) {
    const uint64_t s = static_cast<uint64_t>(k) * P.in_cap + i;
    return s < P.flt_cap ? static_cast<uint32_t>(s) : kInvalid;
}

// Slot of listed candidate q of input i: plane-major like the filtered
// states, so a wave's lanes (consecutive inputs) read and write one plane's
// consecutive entries.
__device__ inline size_t list_slot(const ckf_params& P, uint32_t i, uint32_t q) {
    return static_cast<size_t>(q) * P.in_cap + i;
}

// The link a staged slot becomes (a hole link carries the input's own state).
__device__ inline dlink staged_link(const stage_head& h, const stage_cand& c, uint32_t step,
                                    uint32_t slot) {
    const bool hole = c.meas == kInvalid;
    const uint32_t skip = h.hist & 0xffu, consec = h.hist >> 8;
    dlink L;
    L.prev = h.prev;
    L.seed = h.seed;
    if (hole) {
        L.meas = kInvalid;
        L.packed = pack_link(step, skip + 1u, consec + 1u);
        L.chi2 = FLT_MAX;
        L.chi2_sum = h.chi2_sum;
        L.ndf_sum = h.ndf_sum;
        L.state_ref = h.hole_ref;
    } else {
        L.meas = c.meas & 0x7fffffffu;
        L.packed = pack_link(step, skip, 0u);
        L.chi2 = c.chi2;
        L.chi2_sum = h.chi2_sum + c.chi2;
        L.ndf_sum = h.ndf_sum + ((c.meas >> 31) != 0u ? 2u : 1u);
        L.state_ref = (kRefFiltered << kRefKindShift) | slot;
    }
    return L;
}

// Compaction partitions of the emit kernels: a wave takes kEmitChunks
// consecutive 64-item chunks per ticket, counts them, runs one decoupled
// look-back for the whole partition, then writes.  Large partitions keep the
// look-back chain short (its round trips, not the bytes, bound these kernels).
constexpr uint32_t kEmitChunks = 16;

// Chunks per partition for n items: one per 64 chunks of work, at most
// kEmitChunks -- a look-back round trip covers 64 partitions, so the chain
// any partition waits on stays at a round or two, while small steps still
// spread over many waves.
__device__ inline uint32_t emit_chunks(uint32_t n) {
    const uint32_t c = n / (kWave * 64u);
    return c < 1u ? 1u : (c > kEmitChunks ? kEmitChunks : c);
}

// A listed candidate's MLP row (16 shorts, 32 B): features 0-8, 10 and 11
// normalised and rounded to BF16 exactly as the scorer takes them, short 9
// zero, and feature 9 (standing) as its raw FP32 running value in shorts
// 12-13 -- it needs every seed's minimum of the step, so k_score_update
// finalises it.  Shorts 14-15 are zero.
constexpr int kRowShorts = 16;

__device__ inline void store_list_row(const mlp_dev& W, short* row, const float (&f)[kFeatures]) {
    uint32_t w[8];
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int a = 2 * e, b = 2 * e + 1;
        const uint16_t lo = a == kFeatStanding ? 0u
                                               : static_cast<uint16_t>(f2bf((f[a] - W.mean[a]) * W.inv_std[a]));
        const uint16_t hi = b == kFeatStanding ? 0u
                                               : static_cast<uint16_t>(f2bf((f[b] - W.mean[b]) * W.inv_std[b]));
        w[e] = static_cast<uint32_t>(lo) | (static_cast<uint32_t>(hi) << 16);
    }
    w[6] = __float_as_uint(f[kFeatStanding]);
    w[7] = 0u;
    uint4* dst = reinterpret_cast<uint4*>(row);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// Scorer-pruned finding, first kernel: per input the list_len best candidates
// in (chi2, index) order and their MLP rows, at the input's fixed slots
// list_slot(input, q); k_score_update scores and updates them.
__global__ __launch_bounds__(kBlock) void k_find_list(ckf_params P) {
    step_ctl* ctl = P.ctl;
    if (ctl->done) {
        return;
    }
    const uint32_t step = ctl->step;
    const uint32_t n_in = ctl->n_in < P.in_cap ? ctl->n_in : P.in_cap;
    const uint32_t prev_base = ctl->prev_link_base;
    const uint32_t team = lane() / kListTeam;
    const uint32_t tl = lane() % kListTeam;
    const uint32_t Lc = P.list_len;
    const uint32_t epoch = ctl->epoch;
    const uint32_t n_parts = (n_in + (kWave / kListTeam) - 1) / (kWave / kListTeam);
    // dynamic partitions (dense surfaces make some inputs far heavier)
    xcd_ticket_run queue(&ctl->xcd_ticket[1][0].v, n_parts);
    for (;;) {
        const uint32_t p = queue.next();
        if (p >= n_parts) {
            break;
        }
        const uint32_t i0 = p * (kWave / kListTeam);
        const uint32_t i = i0 + team;
        const bool active = i < n_in;
        find_input in;
        uint32_t n_pass = 0u, k = 0u;
        // candidate q of the list is held by team lane q % kListTeam in slot q / kListTeam
        uint32_t sm[kListSlots];
        float sc[kListSlots];
        if (active) {
            load_input(P, step, i, prev_base, in);
            cand_stream<6, kListTeam> cs;
            stream_init(P, in, tl, cs);
            n_pass = cs.n_pass;
            for (; k < Lc; ++k) {
                float best_c;
                uint32_t best_m;
                stream_next(P, in, tl, cs, best_c, best_m);
                if (best_m == kInvalid) {
                    break;
                }
#pragma unroll
                for (int j = 0; j < kListSlots; ++j) {
                    if (k == static_cast<uint32_t>(j * kListTeam) + tl) {
                        sm[j] = best_m;
                        sc[j] = best_c;
                    }
                }
            }
        }
        if (active) {
            if (tl == 0u) {
                P.list_n[i] = k;
                P.list_seed[i] = in.seed;
            }
            // the team's lanes write the list and the features of their
            // candidates, at the input's fixed slots list_slot(i, q)
#pragma unroll
            for (int j = 0; j < kListSlots; ++j) {
                const uint32_t q = static_cast<uint32_t>(j * kListTeam) + tl;
                if (q < k) {
                    const size_t r = list_slot(P, i, q);
                    P.list_m[r] = sm[j];
                    P.list_c[r] = sc[j];
                    const meas_data md = load_meas(P.meas_pos, P.meas_var, P.meas_info, sm[j]);
                    float f[kFeatures];
                    candidate_features(P, md, in.x, in.c, in.sf, step, in.skip, in.consec,
                                       in.chi2_prev, in.ndf_prev, sc[j], q, n_pass, f);
                    store_list_row(P.scorer, P.list_row + r * kRowShorts, f);
                    if (q == 0u) {
                        standing_offer(P.seed_run_min, in.seed, epoch, f[kFeatStanding]);
                    }
                }
            }
        }
    }
}

// Diagnostics: one scored candidate's verdict as a key a CPU checker replays.
// The key names the candidate by its branch's history, so it does not depend
// on where the branch sits in the step's input order: splitmix64 chained over
// the seed, the step, the measurement of every earlier link of the branch
// (latest first, holes as all-ones) and the candidate's measurement; low 63
// bits, bit 63 = kept.
__device__ inline uint64_t verdict_mix(uint64_t h, uint64_t v) {
    uint64_t z = (h ^ v) + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ inline void log_verdict(const ckf_params& P, uint32_t step, uint32_t input,
                                   uint32_t meas, bool keep) {
    const uint32_t at = atomicAdd(P.verdict_n, 1u);
    if (at >= P.verdict_cap) {
        return;
    }
    uint64_t h = verdict_mix(0u, P.list_seed[input]);
    h = verdict_mix(h, step);
    if (step > 0u) {
        uint32_t g = P.ctl->prev_link_base + P.live[input];
        for (uint32_t k = 0; k < step && g < P.link_cap; ++k) {
            const dlink L = P.links[g];
            h = verdict_mix(h, L.meas);
            if (link_step(L) == 0u) {
                break;
            }
            g = L.prev;
        }
    }
    h = verdict_mix(h, meas);
    P.verdict_log[at] = (h & 0x7fffffffffffffffull) | (keep ? (1ull << 63) : 0ull);
}

// Scorer-pruned finding, second kernel, one input per lane (a wave holds 64
// inputs).  The wave packs its inputs' listed candidates (k_find_list's
// fixed slots) densely into LDS rows -- the features normalised and rounded
// to BF16, feature 9 finalised from its seed's minimum -- kRows at a time,
// scores them 16 rows per matrix-core pass, and keeps one 16-bit verdict mask
// per sub-tile in LDS; then each lane walks its input's kept entries in list
// order, runs the gain update until N succeed and stages the links as the
// plain find does (k_find_emit builds them).
constexpr uint32_t kRows = 128;    // LDS rows of a wave per packing pass
constexpr int kListMax = kListTeam * kListSlots;
constexpr uint32_t kSubTiles = (kWave * kListMax) / 16u;  // sub-tiles of a full partition
static_assert(kListMax <= 32, "kept-entry mask is 32 bits");

// The update of one input after the MLP: its kept entries in list order (a
// mask read from the sub-tiles' verdicts), the gain update until N succeed,
// staging as the plain find.
__device__ __forceinline__ void score_update_input(const ckf_params& P, uint32_t i, uint32_t k,
                                                   uint32_t base, const uint16_t* keep16,
                                                   uint32_t step, uint32_t prev_base) {
    step_ctl* ctl = P.ctl;
    const uint32_t N = P.n_branch_surface;
    uint32_t kept = 0u;
    for (uint32_t q = 0; q < k; ++q) {
        const uint32_t r = base + q;
        if ((keep16[r / 16u] >> (r % 16u)) & 1u) {
            kept |= 1u << q;
        }
    }
    if (P.verdict_log != nullptr) {
        for (uint32_t q = 0; q < k; ++q) {
            log_verdict(P, step, i, P.list_m[list_slot(P, i, q)], ((kept >> q) & 1u) != 0u);
        }
    }
    find_input in;
    load_input<false>(P, step, i, prev_base, in);
    P.stage_in[i] = make_stage_head(in, step, i);
    uint32_t count = 0u;
    while (count < N && kept != 0u) {
        const uint32_t li = static_cast<uint32_t>(__builtin_ctz(kept));
        kept &= kept - 1u;
        const uint32_t best_m = P.list_m[list_slot(P, i, li)];
        const float best_c = P.list_c[list_slot(P, i, li)];
        const meas_data md = load_meas(P.meas_pos, P.meas_var, P.meas_info, best_m);
        float fx[6], fc[21];
#pragma unroll
        for (int e = 0; e < 6; ++e) fx[e] = in.x[e];
#pragma unroll
        for (int e = 0; e < 21; ++e) fc[e] = in.c[e];
        int res = CKF_FIT_SUCCESS;
        if (in.rule0) {
            // step 0 on a surface without material: keep the seed, take the
            // measurement variances
            fc[sym_idx(0, 0)] = md.v0;
            fc[sym_idx(1, 1)] = md.v1;
        } else {
            res = gain_update(md, fx, fc);
        }
        if (res == CKF_FIT_SUCCESS) {
            const uint32_t slot = flt_slot(P, i, count);
            if (slot < P.flt_cap) {
                store_soa(P.flt, P.flt_cap, slot, fx, fc, in.sf);
                P.stage_slot[slot] = make_stage_cand(best_m, best_c, md.dims);
            } else {
                set_overflow(ctl, 4u);
            }
            ++count;
        }
    }
    if (count == 0u) {
        // no candidate kept: the input continues as one hole link
        const uint32_t slot = flt_slot(P, i, 0u);
        if (slot < P.flt_cap) {
            P.stage_slot[slot] = stage_cand{kInvalid, 0.f};
        } else {
            set_overflow(ctl, 4u);
        }
    }
    P.find_count[i] = count;
}

// Three waves per SIMD: the compiler's own allocation lands four registers
// above the three-wave budget, and the hint costs no spill.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8)))
void k_score_update(ckf_params P) {
    __shared__ __attribute__((aligned(16))) short s_rows[kBlock / kWave][kRows * kRowShorts];
    __shared__ uint16_t s_keep[kBlock / kWave][kSubTiles];
    __shared__ mlp_vecs V;
    __shared__ mlp_frags_lds F;
    step_ctl* ctl = P.ctl;
    if (ctl->done) {
        return;  // the whole grid returns: no barrier is left waiting
    }
    mlp_fill_vecs(P.scorer, V, threadIdx.x, blockDim.x);
    mlp_fill_frags(P.scorer, F, threadIdx.x, blockDim.x);
    __syncthreads();
    const uint32_t step = ctl->step;
    const uint32_t n_in = ctl->n_in < P.in_cap ? ctl->n_in : P.in_cap;
    const uint32_t prev_base = ctl->prev_link_base;
    const uint32_t epoch = ctl->epoch;
    const uint32_t l = lane();
    const uint32_t wave = threadIdx.x / kWave;
    short* rows = s_rows[wave];
    uint16_t* keep16 = s_keep[wave];
    const uint32_t n_parts = (n_in + kWave - 1u) / kWave;
    xcd_ticket_run queue(&ctl->xcd_ticket[0][0].v, n_parts);
    for (;;) {
        const uint32_t p = queue.next();
        if (p >= n_parts) {
            break;
        }
        const uint32_t i = p * kWave + l;
        const bool valid = i < n_in;
        const uint32_t k = valid ? P.list_n[i] : 0u;
        const uint32_t seed = valid ? P.list_seed[i] : 0u;
        const uint32_t incl = wave_inclusive_scan(k);
        const uint32_t total = wave_broadcast(incl, kWave - 1);
        const uint32_t base = incl - k;
        // rows [r0, r0 + kRows) per pass: pack, then score 16 at a time
        for (uint32_t r0 = 0u; r0 < total; r0 += kRows) {
            for (uint32_t q = 0; q < k; ++q) {
                const uint32_t r = base + q;
                if (r < r0 || r >= r0 + kRows) {
                    continue;
                }
                const uint4* src =
                    reinterpret_cast<const uint4*>(P.list_row + list_slot(P, i, q) * kRowShorts);
                const uint4 a = src[0];
                uint4 b = src[1];
                // feature 9 (short 9, the high half of b.x) from the raw
                // running value and the seed's minimum; shorts 12-15 zero
                const float st = standing_of(P.seed_run_min, seed, epoch, __uint_as_float(b.z));
                const uint16_t bf9 = static_cast<uint16_t>(mlp_feature_bf(V, st, kFeatStanding));
                b.x = (b.x & 0xffffu) | (static_cast<uint32_t>(bf9) << 16);
                b.z = 0u;
                b.w = 0u;
                uint4* row = reinterpret_cast<uint4*>(rows + (r - r0) * kRowShorts);
                row[0] = a;
                row[1] = b;
            }
            wave_lds_sync();
            const uint32_t n_rows = total - r0 < kRows ? total - r0 : kRows;
            for (uint32_t t = 0; t < (n_rows + 15u) / 16u; ++t) {
                const uint32_t r = t * 16u + (l % 16u);
                bf16x4 xb = {0, 0, 0, 0};
                if (r < n_rows) {
                    xb = load4(rows + r * kRowShorts + 4u * (l / 16u));
                }
                const float s = mlp_logit16_lds(F, V, xb);
                const uint64_t bits = wave_ballot(l < 16u && r < n_rows && s >= P.score_threshold);
                if (l == 0u) {
                    keep16[r0 / 16u + t] = static_cast<uint16_t>(bits & 0xffffull);
                }
            }
            wave_lds_sync();  // the next pass overwrites the rows
        }
        if (l == 0u && total > 0u) {
            atomicAdd(&ctl->n_rows.v, total);
        }
        if (valid) {
            score_update_input(P, i, k, base, keep16, step, prev_base);
        }
    }
}

// A team of kFindTeam lanes handles one input branch.  Plain: the lanes split
// the surface's measurements (the candidate stream) and take candidates in
// (chi2, index) order; logged: the same, with every filtered slot's features
// kept for the branch log.  The gain update runs redundantly on the team (so
// it stays uniform) and the team leader writes.  A wave holds kWave / kT
// inputs.  (The scorer-pruned update is k_score_update.)
enum find_mode : int { kFindPlain = 0, kFindLogged = 1 };

template <int kMode>
__global__ __launch_bounds__(kBlock) void k_find(ckf_params P) {
    constexpr bool logging = (kMode == kFindLogged);
    constexpr int kT = kFindTeam;
    constexpr uint32_t kPerWave = kWave / kT;
    step_ctl* ctl = P.ctl;
    if (ctl->done) {
        return;
    }
    const uint32_t step = ctl->step;
    const uint32_t n_in = ctl->n_in < P.in_cap ? ctl->n_in : P.in_cap;
    const uint32_t prev_base = ctl->prev_link_base;
    const uint32_t n_parts = (n_in + kPerWave - 1) / kPerWave;
    const uint32_t N = P.n_branch_surface;
    const uint32_t team = lane() / kT;
    const uint32_t tl = lane() % kT;
    const bool leader = (tl == 0u);

    xcd_ticket_run queue(&ctl->xcd_ticket[0][0].v, n_parts);
    for (;;) {
        const uint32_t p = queue.next();
        if (p >= n_parts) {
            break;
        }
        const uint32_t i = p * kPerWave + team;
        const bool valid = i < n_in;
        uint32_t count = 0u;
        if (valid) {
            find_input in;
            load_input(P, step, i, prev_base, in);
            if (leader) {
                P.stage_in[i] = make_stage_head(in, step, i);
            }
            uint32_t n_pass = 0u;
            {
                // candidates in stream order; the gain update runs on every
                // lane of the team (uniform), the leader writes
                cand_stream<6, kFindTeam> cs;
                stream_init(P, in, tl, cs);
                n_pass = cs.n_pass;
                bool rank0 = true;
                while (count < N) {
                    float best_c = 0.f;
                    uint32_t best_m = kInvalid;
                    stream_next(P, in, tl, cs, best_c, best_m);
                    if (best_m == kInvalid) {
                        break;
                    }
                    const meas_data md = load_meas(P.meas_pos, P.meas_var, P.meas_info, best_m);
                    if constexpr (logging) {
                        if (rank0 && leader) {
                            standing_offer(P.seed_run_min, in.seed, ctl->epoch,
                                           branch_run(in.chi2_prev, in.ndf_prev, best_c, md.dims));
                        }
                        rank0 = false;
                    }
                    float fx[6], fc[21];
#pragma unroll
                    for (int k = 0; k < 6; ++k) fx[k] = in.x[k];
#pragma unroll
                    for (int k = 0; k < 21; ++k) fc[k] = in.c[k];
                    int res = CKF_FIT_SUCCESS;
                    if (in.rule0) {
                        // step 0 on a surface without material: keep the
                        // seed, take the measurement variances
                        fc[sym_idx(0, 0)] = md.v0;
                        fc[sym_idx(1, 1)] = md.v1;
                    } else {
                        res = gain_update(md, fx, fc);
                    }
                    if (res == CKF_FIT_SUCCESS) {
                        const uint32_t slot = flt_slot(P, i, count);
                        if (leader) {
                            if (slot < P.flt_cap) {
                                store_soa(P.flt, P.flt_cap, slot, fx, fc, in.sf);
                                P.stage_slot[slot] = make_stage_cand(best_m, best_c, md.dims);
                                if constexpr (logging) {
                                    candidate_features(P, md, in.x, in.c, in.sf, step, in.skip, in.consec,
                                                       in.chi2_prev, in.ndf_prev, best_c, count, n_pass,
                                                       P.feat_slot + static_cast<size_t>(slot) * kFeatures);
                                }
                            } else {
                                set_overflow(ctl, 4u);
                            }
                        }
                        ++count;
                    }
                }
            }
            if (count == 0u && leader) {
                // no candidate kept: the input continues as one hole link
                const uint32_t slot = flt_slot(P, i, 0u);
                if (slot < P.flt_cap) {
                    P.stage_slot[slot] = stage_cand{kInvalid, 0.f};
                } else {
                    set_overflow(ctl, 4u);
                }
            }
        }
        if (valid && leader) {
            P.find_count[i] = count;
        }
    }
}

// Links of the step, after k_find: the inputs' link counts (kept branches, or
// one hole link) laid out in input order, then the links written.  Kept apart
// from k_find so a partition with heavy inputs never holds the look-back
// chain of every later partition while it works.
template <bool kLogging>
__global__ __launch_bounds__(kBlock) void k_find_emit(ckf_params P) {
    step_ctl* ctl = P.ctl;
    if (ctl->done) {
        return;
    }
    const uint32_t step = ctl->step;
    const uint32_t n_in = ctl->n_in < P.in_cap ? ctl->n_in : P.in_cap;
    const uint32_t link_base = ctl->link_base;
    const uint32_t epoch = ctl->epoch;
    const uint32_t kc = emit_chunks(n_in);
    const uint32_t part = kWave * kc;
    const uint32_t n_parts = (n_in + part - 1u) / part;
    auto input_seed = [&](uint32_t j) -> uint32_t { return P.stage_in[j].seed; };
    const bool takes = wave_takes_tickets(n_parts);
    for (;;) {
        const uint32_t p = takes ? wave_ticket(&ctl->ticket[4].v) : n_parts;
        if (p >= n_parts) {
            break;
        }
        const uint32_t i0 = p * part;
        uint32_t n_out[kEmitChunks];
        uint32_t total = 0u;
#pragma unroll
        for (int ch = 0; ch < static_cast<int>(kEmitChunks); ++ch) {
            const uint32_t i = i0 + static_cast<uint32_t>(ch) * kWave + lane();
            const uint32_t count = (static_cast<uint32_t>(ch) < kc && i < n_in) ? P.find_count[i] : 0u;
            n_out[ch] = (static_cast<uint32_t>(ch) < kc && i < n_in) ? (count > 0u ? count : 1u) : 0u;
            total += wave_broadcast(wave_inclusive_scan(n_out[ch]), kWave - 1);
        }
        const uint64_t base = wave_lookback(P.lb_lo[0], nullptr, p, epoch,
                                            pack2(total, 0u));
        if (p == n_parts - 1u && lane() == 0u) {
            ctl->n_links = lo32(base) + total;
        }
        uint32_t run = lo32(base);
        uint32_t last_seed = (i0 > 0u && lane() == 0u) ? input_seed(i0 - 1u) : kInvalid;
        last_seed = wave_broadcast(last_seed, 0);
#pragma unroll 1
        for (uint32_t ch = 0; ch < kc; ++ch) {
            if (i0 + ch * kWave >= n_in) {
                break;  // uniform
            }
            const uint32_t i = i0 + ch * kWave + lane();
            const bool valid = i < n_in;
            // select n_out[ch] without a runtime index into the array
            uint32_t no = 0u;
#pragma unroll
            for (int q = 0; q < static_cast<int>(kEmitChunks); ++q) {
                if (static_cast<uint32_t>(q) == ch) no = n_out[q];
            }
            const uint32_t incl = wave_inclusive_scan(no);
            const uint32_t first = run + incl - no;
            run += wave_broadcast(incl, kWave - 1);
            const stage_head h = valid ? P.stage_in[i] : stage_head{};
            const uint32_t seed = valid ? h.seed : kInvalid;
            // head of the seed's contiguous run of inputs
            uint32_t prev_seed = __shfl(seed, (lane() + kWave - 1u) % kWave, kWave);
            if (lane() == 0u) {
                prev_seed = last_seed;
            }
            last_seed = wave_broadcast(seed, kWave - 1);
            if (!valid) {
                continue;
            }
            if (i == 0u || prev_seed != seed) {
                P.seed_first[seed] = first;
            }
            for (uint32_t k = 0; k < no; ++k) {
                const uint32_t slot = flt_slot(P, i, k);
                if (slot >= P.flt_cap) {
                    break;  // k_find raised the overflow that discards this step
                }
                const dlink L = staged_link(h, P.stage_slot[slot], step, slot);
                const uint32_t g = link_base + first + k;
                if (g < P.link_cap && first + k < P.state_cap) {
                    P.links[g] = L;
                    if constexpr (kLogging) {
                        if (L.meas != kInvalid) {
#pragma unroll
                            for (int q = 0; q < kFeatures; ++q) {
                                float v = P.feat_slot[static_cast<size_t>(slot) * kFeatures + q];
                                if (q == kFeatStanding) {
                                    v = standing_of(P.seed_run_min, seed, epoch, v);
                                }
                                P.log_feat[static_cast<size_t>(g) * kFeatures + q] = v;
                            }
                        }
                    }
                } else {
                    set_overflow(ctl, 1u);
                }
            }
        }
    }
}

// Overflow bits raised by k_find (links, filtered states).  Later kernels of
// the step only read these: the bits k_propagate raises itself must not change
// the partition count its waves see mid-kernel.
constexpr uint32_t kFindOverflow = 1u | 4u;

// Links of this step that exist: a step whose find kernel overflowed a buffer
// is abandoned (the host regrows and re-runs), so nothing walks missing links.
__device__ inline uint32_t links_this_step(const ckf_params& P, const step_ctl* ctl) {
    if ((ctl->overflow & kFindOverflow) != 0u) {
        return 0u;
    }
    uint32_t n = ctl->n_links;
    n = n < P.state_cap ? n : P.state_cap;
    const uint32_t room = P.link_cap > ctl->link_base ? P.link_cap - ctl->link_base : 0u;
    return n < room ? n : room;
}

__global__ __launch_bounds__(kBlock) void k_dedup_insert(ckf_params P) {
    const step_ctl* ctl = P.ctl;
    if (ctl->done || ctl->step < P.dup_min_len || (ctl->overflow & kFindOverflow) != 0u) {
        return;
    }
    const uint32_t n = links_this_step(P, ctl);
    const uint32_t base = ctl->link_base;
    const uint32_t epoch = ctl->epoch;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        dlink L = load_link(P, base + i);
        while (L.meas == kInvalid && link_step(L) != 0u) {
            L = load_link(P, L.prev);
        }
        const uint32_t key = (L.meas == kInvalid) ? P.n_meas : L.meas;
        P.alive_prefix[i] = key;
        // list heads carry the step's epoch: a head from an earlier step reads as empty
        const unsigned long long tagged = (static_cast<unsigned long long>(epoch) << 32) | i;
        const unsigned long long old = atomicExch(
            reinterpret_cast<unsigned long long*>(&P.dedup_head[key]), tagged);
        P.dedup_next[i] = (static_cast<uint32_t>(old >> 32) == epoch) ? static_cast<uint32_t>(old)
                                                                       : kInvalid;
    }
}

namespace {
// Is track `a` dominated by `b` (identical measurement history down to a's
// first link)?  traccc host :431-471.
__device__ inline bool history_contained(const ckf_params& P, dlink a, dlink b) {
    for (;;) {
        while (a.meas == kInvalid && link_step(a) != 0u) a = load_link(P, a.prev);
        while (b.meas == kInvalid && link_step(b) != 0u) b = load_link(P, b.prev);
        if (a.meas == b.meas) {
            if (link_step(a) == 0u) return true;
            if (link_step(b) == 0u) return false;
            a = load_link(P, a.prev);
            b = load_link(P, b.prev);
        } else {
            return false;
        }
    }
}
}  // namespace

__global__ __launch_bounds__(kBlock) void k_dedup_mark(ckf_params P) {
    step_ctl* ctl = P.ctl;
    if (ctl->done || ctl->step < P.dup_min_len || (ctl->overflow & kFindOverflow) != 0u) {
        return;
    }
    const uint32_t step = ctl->step;
    const uint32_t n = links_this_step(P, ctl);
    const uint32_t base = ctl->link_base;
    const uint32_t epoch = ctl->epoch;
    const uint32_t n_parts = (n + kWave - 1) / kWave;
    xcd_ticket_run queue(&ctl->xcd_ticket[3][0].v, n_parts, 1u);
    for (;;) {
        const uint32_t p = queue.next();
        if (p >= n_parts) {
            break;
        }
        const uint32_t i = p * kWave + lane();
        const bool valid = i < n;
        bool alive = true;
        uint32_t seed = 0u;
        if (valid) {
            const dlink Li = load_link(P, base + i);
            seed = Li.seed;
            const uint32_t key = P.alive_prefix[i];
            if (step + 1u - link_skipped(Li) > P.dup_min_len && Li.ndf_sum > 5u) {
                const float p_this = prob_chi2(Li.chi2_sum, static_cast<float>(Li.ndf_sum - 5u));
                const uint64_t h = P.dedup_head[key];
                const uint32_t first = (static_cast<uint32_t>(h >> 32) == epoch)
                                           ? static_cast<uint32_t>(h) : kInvalid;
                for (uint32_t j = first; j != kInvalid; j = P.dedup_next[j]) {
                    if (j == i) continue;
                    const dlink Lj = load_link(P, base + j);
                    if (step + 1u - link_skipped(Lj) <= P.dup_min_len || Li.ndf_sum <= 5u) continue;
                    bool dominated = history_contained(P, Li, Lj);
                    const float p_that = prob_chi2(Lj.chi2_sum, static_cast<float>(Lj.ndf_sum - 5u));
                    if (p_this != p_that) {
                        dominated = dominated && (p_that >= p_this);
                    } else {
                        dominated = dominated && (j < i);
                    }
                    if (dominated) {
                        alive = false;
                        break;
                    }
                }
            }
        }
        if (valid) {
            P.alive_prefix[i] = alive ? 0x80000000u : 0u;  // ranks: k_dedup_emit
        }
        (void)seed;
    }
}

// After k_dedup_mark: each link's rank among the step's surviving links
// (alive_prefix = alive bit | rank) and, per seed, the rank of its first link.
__global__ __launch_bounds__(kBlock) void k_dedup_emit(ckf_params P) {
    step_ctl* ctl = P.ctl;
    if (ctl->done || ctl->step < P.dup_min_len || (ctl->overflow & kFindOverflow) != 0u) {
        return;
    }
    const uint32_t n = links_this_step(P, ctl);
    const uint32_t base = ctl->link_base;
    const uint32_t epoch = ctl->epoch;
    const uint32_t kc = emit_chunks(n);
    const uint32_t part = kWave * kc;
    const uint32_t n_parts = (n + part - 1u) / part;
    const bool takes = wave_takes_tickets(n_parts);
    for (;;) {
        const uint32_t p = takes ? wave_ticket(&ctl->ticket[6].v) : n_parts;
        if (p >= n_parts) {
            break;
        }
        const uint32_t i0 = p * part;
        uint64_t am[kEmitChunks];
        uint32_t total = 0u;
#pragma unroll
        for (int ch = 0; ch < static_cast<int>(kEmitChunks); ++ch) {
            const uint32_t i = i0 + static_cast<uint32_t>(ch) * kWave + lane();
            const bool a = static_cast<uint32_t>(ch) < kc && i < n && (P.alive_prefix[i] >> 31) != 0u;
            am[ch] = wave_ballot(a);
            total += static_cast<uint32_t>(__popcll(am[ch]));
        }
        const uint64_t b = wave_lookback(P.lb_lo[1], nullptr, p, epoch,
                                         pack2(total, 0u));
        uint32_t run = lo32(b);
#pragma unroll
        for (int ch = 0; ch < static_cast<int>(kEmitChunks); ++ch) {
            const uint32_t i = i0 + static_cast<uint32_t>(ch) * kWave + lane();
            if (static_cast<uint32_t>(ch) < kc && i < n) {
                const bool a = ((am[ch] >> lane()) & 1ull) != 0ull;
                const uint32_t before = run + prefix_count(am[ch]);
                P.alive_prefix[i] = (a ? 0x80000000u : 0u) | before;
                const uint32_t seed = load_link(P, base + i).seed;
                if (P.seed_first[seed] == i) {
                    P.seed_base[seed] = before;
                }
            }
            run += static_cast<uint32_t>(__popcll(am[ch]));
        }
    }
}

__device__ void advance_step(const ckf_params& P, step_ctl* ctl);

// ---------------------------------------------------------------------------
// propagate_to_next_surface
// ---------------------------------------------------------------------------
// The end of a link's propagation (status from the navigator): on NAV_FOUND
// the covariance (read now; the navigation does not need it, so it is not
// held in registers meanwhile) with the start surface's material at step 0,
// the transport to the new surface, its material, the sanity checks and the
// new state; the outcome flag (bit 0: next input, bits 1-2: tips).
template <int kT>
__device__ inline uint8_t finish_link(const ckf_params& P, const navigator<float, true, kT>& nav,
                                      int status, uint32_t end_sf, uint32_t i, uint32_t ref,
                                      uint32_t step, bool leader) {
    const uint32_t in_par = (step + 1u) & 1u;
    const uint32_t out_par = step & 1u;
    const bool success = (status == NAV_FOUND);
    bool ok = success;
    bool succ = false;
    uint32_t ntip = 0u;
    if (success) {
        float x[6], c[21];
        uint32_t sf = 0u;
        load_ref(P, ref, in_par, x, c, sf);
        if (step == 0u) {
            add_material(P, sf, x, c);
        }
        const geo_view G{P.sfs, P.n_layers, P.field};
        finish_transport<kT>(G, P.nav, nav, end_sf, x, c);
        sf = end_sf;
        add_material(P, sf, x, c);
        const float th = x[e_theta];
        if (th <= 0.f || th >= 6.28318530717958647692f) ok = false;
        if (!isfinite(x[e_phi])) ok = false;
        if (fabsf(x[e_qop]) == 0.f) ok = false;
        if (ok) {
            if (leader) {
                store_soa(P.state[out_par], P.state_cap, i, x, c, sf);
            }
            succ = true;
        }
    }
    if (!ok && step + 1u >= P.min_cands) ++ntip;
    if (success && step == P.max_cands - 1u) ++ntip;
    return static_cast<uint8_t>((succ ? 1u : 0u) | (ntip << 1));
}

// kT lanes propagate one link together (team navigator); 64 / kT links per
// wave partition.  The team leader writes the results.  With one lane per
// link, a propagation that needs a second layer search (a few % of them, but
// nearly every wave has one) is written to the defer buffer and continued by
// k_propagate_resume, so the wave does not run a second search, approach and
// window for one or two of its lanes.
template <int kT>
__global__ __launch_bounds__(kBlock) void k_propagate(ckf_params P) {
    step_ctl* ctl = P.ctl;
    if (ctl->done) {
        return;
    }
    // the navigation layers and groups live in LDS for the whole kernel
    __shared__ nav_tables s_nav;
    stage_nav_tables(s_nav, P.layers, P.n_layers, P.groups, P.n_groups, P.nav.mask_tol_max);
    const uint32_t step = ctl->step;
    const uint32_t n = links_this_step(P, ctl);
    const uint32_t base = ctl->link_base;
    constexpr uint32_t kLinksPerWave = kWave / kT;
    const uint32_t n_parts = (n + kLinksPerWave - 1) / kLinksPerWave;
    const uint32_t in_par = (step + 1u) & 1u;
    const bool dedup = step >= P.dup_min_len;
    const uint32_t tl = lane() % kT;
    const bool leader = (tl == 0u);
    const bool defer = (kT == 1) && P.defer_cap > 0u;

    xcd_ticket_run queue(&ctl->xcd_ticket[2][0].v, n_parts, 1u);
    for (;;) {
        const uint32_t p = queue.next();
        if (p >= n_parts) {
            break;
        }
        const uint32_t i = p * kLinksPerWave + lane() / kT;
        const bool valid = i < n;
        bool propagated = false;
        bool deferred = false;
        uint8_t flag = 0u;
        if (valid) {
            const dlink L = load_link(P, base + i);
            bool alive = true;
            uint32_t rank = 0u;
            if (dedup) {
                const uint32_t ap = P.alive_prefix[i];
                alive = (ap >> 31) != 0u;
                rank = (ap & 0x7fffffffu) - P.seed_base[L.seed];
            } else {
                rank = i - P.seed_first[L.seed];
            }
            if (alive && rank < P.n_branch_seed) {
                if (link_skipped(L) > P.max_skip || link_consec(L) > P.max_consec) {
                    flag = static_cast<uint8_t>(1u << 1);
                } else {
                    float x[6];
                    uint32_t sf = 0u;
                    {
                        float c_later[21];
                        load_ref(P, L.state_ref, in_par, x, c_later, sf);
                    }
                    propagated = true;
                    const geo_view G{P.sfs, P.n_layers, P.field};
                    navigator<float, true, kT> nav =
                        make_navigator<kT>(G, s_nav.layers, P.nav, tl, &s_nav, P.n_groups);
                    bound_to_free<float, true>(P.sfs[sf], x, nav.st, nav.J);
                    uint32_t end_sf = 0u;
                    int status = nav.propagate(static_cast<int>(sf), end_sf, defer);
                    if (status == NAV_DEFER) {
                        const uint32_t k = atomicAdd(&ctl->n_defer.v, 1u);
                        if (k < P.defer_cap) {
                            uint32_t w[kDeferWords];
                            w[0] = i;
                            w[1] = static_cast<uint32_t>(nav.deferred_visited);
                            w[2] = static_cast<uint32_t>(nav.deferred_visited >> 32);
                            w[3] = nav.n_steps;
                            w[4] = __float_as_uint(nav.path);
                            const float fs[11] = {nav.st.pos.x, nav.st.pos.y, nav.st.pos.z,
                                                  nav.st.dir.x, nav.st.dir.y, nav.st.dir.z,
                                                  nav.st.qop,   nav.st.time,  nav.st.lo.x,
                                                  nav.st.lo.y,  nav.st.lo.z};
#pragma unroll
                            for (int e = 0; e < 11; ++e) w[5 + e] = __float_as_uint(fs[e]);
#pragma unroll
                            for (int r = 0; r < 3; ++r) {
#pragma unroll
                                for (int j = 0; j < 5; ++j) {
                                    w[16 + 5 * r + j] = __float_as_uint(nav.J.p[r][j]);
                                    w[31 + 5 * r + j] = __float_as_uint(nav.J.d[r][j]);
                                }
                            }
                            w[46] = __float_as_uint(nav.J.t_qop);
                            w[47] = L.state_ref;
#pragma unroll
                            for (uint32_t e = 0; e < kDeferWords; ++e) {
                                P.defer[static_cast<size_t>(e) * P.defer_cap + k] = w[e];
                            }
                            deferred = true;
                        } else {
                            // no room: continue here
                            status = nav.layer_loop(nav.deferred_visited, 1u, end_sf, false);
                        }
                    }
                    if (!deferred) {
                        flag = finish_link<kT>(P, nav, status, end_sf, i, L.state_ref, step, leader);
                    }
                }
            }
        }
        // the link's outcome for k_prop_emit (a deferred one's comes from
        // k_propagate_resume); the propagation count here
        if (valid && leader && !deferred) {
            P.prop_flag[i] = flag;
        }
        const uint64_t mp = wave_ballot(leader && propagated);
        if (lane() == 0u && mp != 0ull) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->n_props.v),
                      static_cast<unsigned long long>(__popcll(mp)));
        }
    }
}

// The deferred propagations of k_propagate<1>, one lane each: the layer
// loop continues from its second iteration with the saved state, then the
// same end as in k_propagate.
__global__ __launch_bounds__(kBlock) void k_propagate_resume(ckf_params P) {
    step_ctl* ctl = P.ctl;
    if (ctl->done) {
        return;
    }
    __shared__ nav_tables s_nav;
    stage_nav_tables(s_nav, P.layers, P.n_layers, P.groups, P.n_groups, P.nav.mask_tol_max);
    const uint32_t step = ctl->step;
    const uint32_t nd = ctl->n_defer.v;
    const uint32_t n = nd < P.defer_cap ? nd : P.defer_cap;
    const uint32_t n_parts = (n + kWave - 1u) / kWave;
    const bool takes = wave_takes_tickets(n_parts);
    for (;;) {
        const uint32_t p = takes ? wave_ticket(&ctl->ticket[7].v) : n_parts;
        if (p >= n_parts) {
            break;
        }
        const uint32_t k = p * kWave + lane();
        if (k >= n) {
            continue;
        }
        uint32_t w[kDeferWords];
#pragma unroll
        for (uint32_t e = 0; e < kDeferWords; ++e) {
            w[e] = P.defer[static_cast<size_t>(e) * P.defer_cap + k];
        }
        const geo_view G{P.sfs, P.n_layers, P.field};
        navigator<float, true, 1> nav = make_navigator<1>(G, s_nav.layers, P.nav, 0u, &s_nav, P.n_groups);
        const uint64_t visited = static_cast<uint64_t>(w[1]) | (static_cast<uint64_t>(w[2]) << 32);
        nav.n_steps = w[3];
        nav.path = __uint_as_float(w[4]);
        nav.st.pos = {__uint_as_float(w[5]), __uint_as_float(w[6]), __uint_as_float(w[7])};
        nav.st.dir = {__uint_as_float(w[8]), __uint_as_float(w[9]), __uint_as_float(w[10])};
        nav.st.qop = __uint_as_float(w[11]);
        nav.st.time = __uint_as_float(w[12]);
        nav.st.lo = {__uint_as_float(w[13]), __uint_as_float(w[14]), __uint_as_float(w[15])};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                nav.J.p[r][j] = __uint_as_float(w[16 + 5 * r + j]);
                nav.J.d[r][j] = __uint_as_float(w[31 + 5 * r + j]);
            }
        }
        nav.J.t_qop = __uint_as_float(w[46]);
        uint32_t end_sf = 0u;
        const int status = nav.layer_loop(visited, 1u, end_sf, false);
        const uint32_t i = w[0];
        P.prop_flag[i] = finish_link<1>(P, nav, status, end_sf, i, w[47], step, true);
    }
}

// After k_propagate: the next step's inputs (links that reached a surface)
// and the tips, compacted in link order (partitions of up to 1024 links); the
// last block to leave runs the step advance (fused_advance) so the step loop
// never waits on the host.
__global__ __launch_bounds__(kBlock) void k_prop_emit(ckf_params P, int fused_advance) {
    __shared__ uint32_t s_exit;  // waves of this block past the queue
    step_ctl* ctl = P.ctl;
    if (ctl->done) {
        return;  // the whole grid returns
    }
    if (threadIdx.x == 0u) {
        s_exit = 0u;
    }
    __syncthreads();
    const uint32_t n = links_this_step(P, ctl);
    const uint32_t base = ctl->link_base;
    const uint32_t epoch = ctl->epoch;
    const uint32_t tip_base = ctl->n_tips;
    const uint32_t kc = emit_chunks(n);
    const uint32_t part = kWave * kc;
    const uint32_t n_parts = (n + part - 1u) / part;
    const bool takes = wave_takes_tickets(n_parts);
    for (;;) {
        const uint32_t p = takes ? wave_ticket(&ctl->ticket[5].v) : n_parts;
        if (p >= n_parts) {
            break;
        }
        const uint32_t i0 = p * part;
        uint32_t fl[kEmitChunks];
        uint32_t s_tot = 0u, t_tot = 0u;
#pragma unroll
        for (int ch = 0; ch < static_cast<int>(kEmitChunks); ++ch) {
            const uint32_t i = i0 + static_cast<uint32_t>(ch) * kWave + lane();
            fl[ch] = (static_cast<uint32_t>(ch) < kc && i < n) ? P.prop_flag[i] : 0u;
            s_tot += static_cast<uint32_t>(__popcll(wave_ballot((fl[ch] & 1u) != 0u)));
            t_tot += static_cast<uint32_t>(__popcll(wave_ballot((fl[ch] >> 1) >= 1u)) +
                                           __popcll(wave_ballot((fl[ch] >> 1) >= 2u)));
        }
        const uint64_t b = wave_lookback(P.lb_lo[2], P.lb_hi[2], p, epoch,
                                         pack2(s_tot, t_tot));
        uint32_t s_run = lo32(b), t_run = hi32(b);
#pragma unroll
        for (int ch = 0; ch < static_cast<int>(kEmitChunks); ++ch) {
            const uint32_t i = i0 + static_cast<uint32_t>(ch) * kWave + lane();
            const uint32_t f = fl[ch];
            const bool succ = (f & 1u) != 0u;
            const uint32_t ntip = f >> 1;
            const uint64_t ms = wave_ballot(succ);
            const uint64_t m1 = wave_ballot(ntip >= 1u);
            const uint64_t m2 = wave_ballot(ntip >= 2u);
            if (succ) {
                const uint32_t pos = s_run + prefix_count(ms);
                if (pos < P.in_cap) {
                    P.live[pos] = i;
                } else {
                    set_overflow(ctl, 16u);
                }
            }
            if (ntip > 0u) {
                uint32_t pos = tip_base + t_run + prefix_count(m1) + prefix_count(m2);
                for (uint32_t k = 0; k < ntip; ++k, ++pos) {
                    if (pos < P.tip_cap) {
                        P.tips[pos] = base + i;
                    } else {
                        set_overflow(ctl, 8u);
                    }
                }
            }
            s_run += static_cast<uint32_t>(__popcll(ms));
            t_run += static_cast<uint32_t>(__popcll(m1) + __popcll(m2));
        }
        if (p == n_parts - 1u && lane() == 0u && !fused_advance) {
            ctl->n_live = lo32(b) + s_tot;
            ctl->n_tips_step = hi32(b) + t_tot;
        }
    }
    if (fused_advance) {
        // The last block to leave runs the step advance.  Every wave of the
        // grid (ticket holder or not) counts itself once it has read the
        // control block -- in LDS, then the block's last wave with one device
        // atomic -- so no wave can start on the advanced control block.  The
        // step's totals come from the last partition's look-back words, which
        // carry their own visibility; the overflow bits that other blocks of
        // this kernel raise are read with a device-scope atomic load (the
        // bits they set with atomicOr live at the device's coherence point);
        // everything else the advance reads was written by earlier kernels,
        // and its own stores reach the next kernel at the launch boundary.
        uint32_t in_block = 0u;
        if (lane() == 0u) {
            in_block = atomicAdd(&s_exit, 1u) + 1u;
        }
        in_block = wave_broadcast(in_block, 0);
        if (in_block == blockDim.x / kWave) {
            uint32_t before = 0u;
            if (lane() == 0u) {
                before = atomicAdd(&ctl->exit_count.v, 1u);
            }
            before = wave_broadcast(before, 0);
            if (before + 1u == gridDim.x) {
                uint32_t live = 0u, tips = 0u;
                if (n_parts > 0u) {
                    live = lookback_inclusive(P.lb_lo[2], n_parts - 1u, epoch);
                    tips = lookback_inclusive(P.lb_hi[2], n_parts - 1u, epoch);
                }
                if (lane() == 0u) {
                    ctl->n_live = live;
                    ctl->n_tips_step = tips;
                    advance_step(P, ctl);
                    ctl->exit_count.v = 0u;
                }
            }
        }
    }
}

// Step advance: move the control block to the next step.  Run by one lane
// once the step's propagation and its compaction have finished (k_advance,
// or the last block to leave k_prop_emit).
__device__ void advance_step(const ckf_params& P, step_ctl* ctl) {
    const uint32_t n_links = ctl->n_links;
    ctl->n_links_total += n_links;
    ctl->n_tips += ctl->n_tips_step;
    ctl->n_tips_step = 0u;
    ctl->prev_link_base = ctl->link_base;
    ctl->link_base += n_links;
    ctl->n_in = ctl->n_live;
    ctl->n_live = 0u;
    ctl->n_links = 0u;
    ctl->n_scored += ctl->n_rows.v;
    ctl->n_rows.v = 0u;
    ctl->step += 1u;
    ctl->n_steps_run = ctl->step;
    ctl->epoch += 1u;
    for (int k = 0; k < kTickets; ++k) ctl->ticket[k].v = 0u;
    ctl->n_defer.v = 0u;
    for (int q = 0; q < 4; ++q) {
        for (uint32_t x = 0; x < kXcds; ++x) ctl->xcd_ticket[q][x].v = 0u;
    }
    const uint32_t overflow =
        __hip_atomic_load(&ctl->overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (ctl->n_in == 0u || ctl->step >= P.max_cands || overflow != 0u) {
        ctl->done = 1u;
    }
}

__global__ void k_advance(ckf_params P) {
    step_ctl* ctl = P.ctl;
    if (ctl->done) {
        return;
    }
    advance_step(P, ctl);
}

// ---------------------------------------------------------------------------
// build_tracks
// ---------------------------------------------------------------------------
// Tips in (step, event) segments: each tip's kept rank and constituent
// prefix, and every segment's bounds.  A partition is emit_chunks(n) chunks
// of 64 tips with one look-back, so the look-back chain stays short when
// there are millions of tips; the (event, step) key of a tip is
// ev * max_cands + step (the host keeps n_events * max_cands below 2^31).
__global__ __launch_bounds__(kBlock) void k_build_scan(ckf_params P) {
    step_ctl* ctl = P.ctl;
    const uint32_t n = ctl->n_tips < P.tip_cap ? ctl->n_tips : P.tip_cap;
    const uint32_t epoch = ctl->epoch;
    const uint32_t kc = emit_chunks(n);
    const uint32_t part = kWave * kc;
    const uint32_t n_parts = (n + part - 1u) / part;
    const uint32_t S = P.max_cands;
    auto seg_key = [&](uint32_t g) -> uint32_t {
        const dlink L = load_link(P, g);
        return P.seed_event[L.seed] * S + link_step(L);
    };
    const bool takes = wave_takes_tickets(n_parts);
    for (;;) {
        const uint32_t p = takes ? wave_ticket(&ctl->ticket[3].v) : n_parts;
        if (p >= n_parts) {
            break;
        }
        const uint32_t t0 = p * part;
        uint32_t kn[kEmitChunks];   // keep << 31 | n_cands
        uint32_t key[kEmitChunks];  // segment key
        uint32_t k_tot = 0u, c_tot = 0u;
#pragma unroll
        for (int ch = 0; ch < static_cast<int>(kEmitChunks); ++ch) {
            const uint32_t t = t0 + static_cast<uint32_t>(ch) * kWave + lane();
            kn[ch] = 0u;
            key[ch] = 0u;
            if (static_cast<uint32_t>(ch) < kc && t < n) {
                const dlink L = load_link(P, P.tips[t]);
                const uint32_t st = link_step(L);
                const uint32_t nc = st + 1u - link_skipped(L);
                const bool keep = nc >= P.min_cands && nc <= P.max_cands;
                kn[ch] = (keep ? 0x80000000u : 0u) | nc;
                key[ch] = P.seed_event[L.seed] * S + st;
            }
            const bool keep = (kn[ch] >> 31) != 0u;
            k_tot += static_cast<uint32_t>(__popcll(wave_ballot(keep)));
            c_tot += wave_broadcast(wave_inclusive_scan(keep ? (kn[ch] & 0x7fffffffu) : 0u), kWave - 1);
        }
        const uint64_t b = wave_lookback(P.lb_lo[3], P.lb_hi[3], p, epoch,
                                         pack2(k_tot, c_tot));
        uint32_t k_run = lo32(b), c_run = hi32(b);
        // segment key of the tip before this partition's first
        uint32_t key_before = 0u;
        if (t0 > 0u && t0 < n) {
            key_before = seg_key(P.tips[t0 - 1u]);
        }
#pragma unroll
        for (int ch = 0; ch < static_cast<int>(kEmitChunks); ++ch) {
            if (static_cast<uint32_t>(ch) >= kc) {
                continue;  // kc is wave-uniform
            }
            const uint32_t t = t0 + static_cast<uint32_t>(ch) * kWave + lane();
            const bool keep = (kn[ch] >> 31) != 0u;
            const uint32_t nc = kn[ch] & 0x7fffffffu;
            const uint32_t w = keep ? nc : 0u;
            const uint32_t cincl = wave_inclusive_scan(w);
            const uint64_t mk = wave_ballot(keep);
            // the previous tip's key: the lane below, across chunk boundaries
            const uint32_t below = __shfl(key[ch], static_cast<int>((lane() + kWave - 1u) % kWave), kWave);
            const uint32_t prev_key = lane() == 0u ? key_before : below;
            if (t < n) {
                const uint32_t kp = k_run + prefix_count(mk);
                const uint32_t cp = c_run + cincl - w;
                P.tip_keep[t] = kn[ch];
                P.tip_rank[t] = kp;
                P.tip_cprefix[t] = cp;
                const uint32_t seg = key[ch];
                if (t == 0u || prev_key != seg) {
                    P.seg_base[seg] = kp;
                    P.seg_cbase[seg] = cp;
                    // the previous segment ends where this one starts
                    if (t > 0u) {
                        P.seg_count[prev_key] = kp;
                        P.seg_ccount[prev_key] = cp;
                    }
                }
                if (t + 1u == n) {
                    P.seg_count[seg] = kp + (keep ? 1u : 0u);
                    P.seg_ccount[seg] = cp + w;
                }
            }
            key_before = wave_broadcast(key[ch], kWave - 1);
            k_run += static_cast<uint32_t>(__popcll(mk));
            c_run += wave_broadcast(cincl, kWave - 1);
        }
    }
}

// One workgroup: exclusive scan of the (event, step) table, event offsets and
// totals.
__global__ __launch_bounds__(1024) void k_build_table(ckf_params P) {
    __shared__ uint32_t s_a[1024];
    __shared__ uint32_t s_b[1024];
    const uint32_t S = P.max_cands;
    const uint32_t total = P.n_events * S;
    const uint32_t per = (total + blockDim.x - 1) / blockDim.x;
    const uint32_t lo = threadIdx.x * per;
    const uint32_t hi = lo + per < total ? lo + per : total;
    // seg_count / seg_ccount hold the kept / constituent prefix at the end of
    // each present segment (0 for absent ones); turn them into counts
    for (uint32_t k = lo; k < hi; ++k) {
        const uint32_t e = P.seg_count[k];
        const uint32_t ce = P.seg_ccount[k];
        P.seg_count[k] = e - (e ? P.seg_base[k] : 0u);
        P.seg_ccount[k] = ce - (e || ce ? P.seg_cbase[k] : 0u);
    }
    uint32_t a = 0u, b = 0u;
    for (uint32_t k = lo; k < hi; ++k) {
        a += P.seg_count[k];
        b += P.seg_ccount[k];
    }
    s_a[threadIdx.x] = a;
    s_b[threadIdx.x] = b;
    __syncthreads();
    for (uint32_t d = 1; d < blockDim.x; d <<= 1) {
        uint32_t va = 0u, vb = 0u;
        if (threadIdx.x >= d) {
            va = s_a[threadIdx.x - d];
            vb = s_b[threadIdx.x - d];
        }
        __syncthreads();
        s_a[threadIdx.x] += va;
        s_b[threadIdx.x] += vb;
        __syncthreads();
    }
    uint32_t ra = s_a[threadIdx.x] - a;
    uint32_t rb = s_b[threadIdx.x] - b;
    for (uint32_t k = lo; k < hi; ++k) {
        if (k % S == 0u) {
            P.event_track_off[k / S] = ra;
            P.event_const_off[k / S] = rb;
        }
        P.seg_off[k] = ra;
        P.seg_coff[k] = rb;
        ra += P.seg_count[k];
        rb += P.seg_ccount[k];
    }
    if (threadIdx.x == blockDim.x - 1) {
        P.event_track_off[P.n_events] = s_a[threadIdx.x];
        P.event_const_off[P.n_events] = s_b[threadIdx.x];
        P.ctl->n_tracks = s_a[threadIdx.x];
        P.ctl->n_constituents = s_b[threadIdx.x];
    }
}

// Ranking key of a kept tip among its seed's tips: more measurements first,
// then the lower chi2 sum, then the lower tip index.
__device__ inline uint64_t tip_rank_key(uint32_t n_cands, float chi2_sum, uint32_t t) {
    const uint32_t cb = __float_as_uint(chi2_sum >= 0.f ? chi2_sum : 0.f);  // order-preserving
    return (static_cast<uint64_t>(n_cands & 0xffu) << 56) |
           (static_cast<uint64_t>(~cb) << 24) |
           static_cast<uint64_t>((0xffffffu - (t < 0xffffffu ? t : 0xffffffu)) & 0xffffffu);
}

// Branch log: each seed's best kept tip (its survive labels go to that
// track's links only).
__global__ __launch_bounds__(kBlock) void k_label_best(ckf_params P) {
    const step_ctl* ctl = P.ctl;
    const uint32_t n = ctl->n_tips < P.tip_cap ? ctl->n_tips : P.tip_cap;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t kv = P.tip_keep[t];
        if ((kv >> 31) == 0u) {
            continue;
        }
        const dlink L = load_link(P, P.tips[t]);
        atomicMax(reinterpret_cast<unsigned long long*>(&P.seed_best[L.seed]),
                  static_cast<unsigned long long>(tip_rank_key(kv & 0x7fffffffu, L.chi2_sum, t)));
    }
}

__global__ __launch_bounds__(kBlock) void k_build_write(ckf_params P) {
    const step_ctl* ctl = P.ctl;
    const uint32_t n = ctl->n_tips < P.tip_cap ? ctl->n_tips : P.tip_cap;
    const uint32_t S = P.max_cands;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t kv = P.tip_keep[t];
        if ((kv >> 31) == 0u) {
            continue;
        }
        const uint32_t nc = kv & 0x7fffffffu;
        uint32_t g = P.tips[t];
        dlink L = load_link(P, g);
        const uint32_t ev = P.seed_event[L.seed];
        const size_t seg = static_cast<size_t>(ev) * S + link_step(L);
        const uint32_t pos = P.seg_off[seg] + (P.tip_rank[t] - P.seg_base[seg]);
        const uint32_t cpos = P.seg_coff[seg] + (P.tip_cprefix[t] - P.seg_cbase[seg]);
        if (pos >= P.out_cap || cpos + nc > P.out_const_cap) {
            continue;
        }
        const uint32_t moff = P.event_meas_off[ev];
        // logged runs: only the seed's best track marks its links as survivors
        const bool mark = P.log_label != nullptr &&
                          P.seed_best[L.seed] == tip_rank_key(nc, L.chi2_sum, t);
        if (P.vote_key != nullptr) {
            // p-value of the tip's running sums ranks the track in measurement votes
            const float vp = prob_chi2(L.chi2_sum, static_cast<float>(L.ndf_sum) - 5.f);
            P.vote_key[pos] = (static_cast<uint64_t>(__float_as_uint(vp)) << 32) |
                              static_cast<uint64_t>(0xffffffffu - pos);
        }
        float ndf = 0.f, chi2 = 0.f;
        for (int k = static_cast<int>(nc) - 1; k >= 0; --k) {
            while (L.meas == kInvalid && link_step(L) != 0u) {
                g = L.prev;
                L = load_link(P, g);
            }
            P.out_const[cpos + static_cast<uint32_t>(k)] = L.meas - moff;
            ndf += static_cast<float>(P.meas_info[L.meas] & 0xffu);
            chi2 += L.chi2;
            if (mark) {
                P.log_label[g] = 1u;  // this measurement link is on the seed's best track
            }
            if (k > 0) {
                g = L.prev;
                L = load_link(P, g);
            }
        }
        ndf -= 5.f;
        const uint32_t s = L.seed;
        P.out_outcome[pos] = 0u;
        P.out_surface[pos] = P.seed_surface[s];
        copy_track_params(P.seed_vec + 6u * static_cast<size_t>(s), P.seed_cov + 36u * static_cast<size_t>(s),
                          P.out_vec + 6u * static_cast<size_t>(pos), P.out_cov + 36u * static_cast<size_t>(pos));
        P.out_ndf[pos] = ndf;
        P.out_chi2[pos] = chi2;
        P.out_pval[pos] = prob_chi2(chi2, ndf);
        P.out_nholes[pos] = link_skipped(L);
        P.out_offsets[pos] = cpos;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.out_offsets[ctl->n_tracks < P.out_cap ? ctl->n_tracks : P.out_cap] = ctl->n_constituents;
    }
}

// ---------------------------------------------------------------------------
// diagnostics: single pieces of the step pipeline on plain arrays
// ---------------------------------------------------------------------------
__global__ void k_debug_kf(uint32_t n, const float* vec, const float* cov, const float2* meas,
                           const float2* var, const uint32_t* info, float* chi2, int32_t* status,
                           float* out_vec, float* out_cov) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    float x[6], c[21];
    #pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = vec[6u * i + k];
    #pragma unroll
    for (int r = 0; r < 6; ++r)
        #pragma unroll
        for (int q = r; q < 6; ++q) c[sym_idx(r, q)] = cov[36u * i + 6u * r + q];
    const meas_data md = load_meas(meas, var, info, i);
    chi2[i] = predicted_chi2(md, x, c);
    const int st = gain_update(md, x, c);
    status[i] = st;
    #pragma unroll
    for (int k = 0; k < 6; ++k) out_vec[6u * i + k] = x[k];
    #pragma unroll
    for (int r = 0; r < 6; ++r)
        #pragma unroll
        for (int q = 0; q < 6; ++q) out_cov[36u * i + 6u * r + q] = c[sym_idx(r, q)];
}

__global__ void k_debug_propagate(ckf_params P, uint32_t n, const uint32_t* start, const float* vec,
                                  const float* cov, int apply_start_material, int32_t* status,
                                  uint32_t* end_surface, float* out_vec, float* out_cov, float* path,
                                  uint32_t* n_steps) {
    __shared__ nav_tables s_nav;
    stage_nav_tables(s_nav, P.layers, P.n_layers, P.groups, P.n_groups, P.nav.mask_tol_max);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    float x[6], c[21];
    #pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = vec[6u * i + k];
    #pragma unroll
    for (int r = 0; r < 6; ++r)
        #pragma unroll
        for (int q = r; q < 6; ++q) c[sym_idx(r, q)] = cov[36u * i + 6u * r + q];
    const uint32_t sf = start[i];
    if (apply_start_material) {
        add_material(P, sf, x, c);
    }
    navigator<float, true> nav;
    nav.sfs = P.sfs;
    nav.layers = s_nav.layers;
    nav.n_layers = P.n_layers;
    nav.groups = s_nav.groups;
    nav.n_groups = P.n_groups;
    nav.layer_boxes = s_nav.boxes;
    nav.group_boxes = s_nav.boxes + 64;
    nav.field = P.field;
    nav.cfg = P.nav;
    bound_to_free<float, true>(P.sfs[sf], x, nav.st, nav.J);
    uint32_t end_sf = 0u;
    const int st = nav.propagate(static_cast<int>(sf), end_sf);
    status[i] = st;
    path[i] = nav.path;
    n_steps[i] = nav.n_steps;
    end_surface[i] = end_sf;
    if (st != NAV_FOUND) {
        return;
    }
    float nx[6];
    float Jb[6][6];
    free_to_bound<float, true>(P.sfs[end_sf], nav.st, nav.J, P.field, P.nav.mass, P.nav.abs_q, nx,
                               Jb);
    float M[6][6];
    #pragma unroll
    for (int r = 0; r < 6; ++r)
        #pragma unroll
        for (int q = 0; q < 6; ++q) {
            float acc = 0.f;
            #pragma unroll
            for (int k = 0; k < 6; ++k) acc += Jb[r][k] * c[sym_idx(k, q)];
            M[r][q] = acc;
        }
    float nc[21];
    #pragma unroll
    for (int r = 0; r < 6; ++r)
        #pragma unroll
        for (int q = r; q < 6; ++q) {
            float acc = 0.f;
            #pragma unroll
            for (int k = 0; k < 6; ++k) acc += M[r][k] * Jb[q][k];
            nc[sym_idx(r, q)] = acc;
        }
    add_material(P, end_sf, nx, nc);
    #pragma unroll
    for (int k = 0; k < 6; ++k) out_vec[6u * i + k] = nx[k];
    #pragma unroll
    for (int r = 0; r < 6; ++r)
        #pragma unroll
        for (int q = 0; q < 6; ++q) out_cov[36u * i + 6u * r + q] = nc[sym_idx(r, q)];
}

hipError_t launch_debug_kf(uint32_t n, const float* vec, const float* cov, const float2* meas,
                           const float2* var, const uint32_t* info, float* chi2, int32_t* status,
                           float* out_vec, float* out_cov, hipStream_t stream) {
    hipLaunchKernelGGL(k_debug_kf, dim3((n + 255u) / 256u), dim3(256), 0, stream, n, vec, cov, meas,
                       var, info, chi2, status, out_vec, out_cov);
    return hipGetLastError();
}

hipError_t launch_debug_propagate(const ckf_params& P, uint32_t n, const uint32_t* start,
                                  const float* vec, const float* cov, int apply_start_material,
                                  int32_t* status, uint32_t* end_surface, float* out_vec,
                                  float* out_cov, float* path, uint32_t* n_steps,
                                  hipStream_t stream) {
    hipLaunchKernelGGL(k_debug_propagate, dim3((n + 63u) / 64u), dim3(64), 0, stream, P, n, start,
                       vec, cov, apply_start_material, status, end_surface, out_vec, out_cov, path,
                       n_steps);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_find_work(const ckf_params& P, const launch_shape& s, hipStream_t stream) {
    if (P.scored) {
        hipLaunchKernelGGL(k_find_list, dim3(s.list_blocks), dim3(kBlock), 0, stream, P);
        hipLaunchKernelGGL(k_score_update, dim3(s.score_blocks), dim3(kBlock), 0, stream, P);
    } else if (P.feat_slot != nullptr) {
        hipLaunchKernelGGL(k_find<kFindLogged>, dim3(s.find_blocks), dim3(kBlock), 0, stream, P);
    } else {
        hipLaunchKernelGGL(k_find<kFindPlain>, dim3(s.find_blocks), dim3(kBlock), 0, stream, P);
    }
    return hipGetLastError();
}

hipError_t launch_find_emit(const ckf_params& P, const launch_shape& s, hipStream_t stream) {
    if (P.feat_slot != nullptr) {
        hipLaunchKernelGGL(k_find_emit<true>, dim3(s.emit_blocks), dim3(kBlock), 0, stream, P);
    } else {
        hipLaunchKernelGGL(k_find_emit<false>, dim3(s.emit_blocks), dim3(kBlock), 0, stream, P);
    }
    return hipGetLastError();
}

hipError_t launch_find(const ckf_params& P, const launch_shape& s, hipStream_t stream) {
    const hipError_t e = launch_find_work(P, s, stream);
    if (e != hipSuccess) {
        return e;
    }
    return launch_find_emit(P, s, stream);
}

// Branch log compaction: measurement links [0, n) in order, 1024 per block.
__global__ __launch_bounds__(1024) void k_log_count(ckf_params P, uint32_t n, uint32_t* block_n) {
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    const uint32_t g = blockIdx.x * 1024u + threadIdx.x;
    if (g < n && P.links[g].meas != kInvalid) atomicAdd(&s_n, 1u);
    __syncthreads();
    if (threadIdx.x == 0) block_n[blockIdx.x] = s_n;
}

__global__ __launch_bounds__(1024) void k_log_write(ckf_params P, uint32_t n, const uint32_t* block_off,
                                                    float* features, uint8_t* labels, uint32_t cap) {
    __shared__ uint32_t s_w[16];
    const uint32_t g = blockIdx.x * 1024u + threadIdx.x;
    const bool is_meas = g < n && P.links[g].meas != kInvalid;
    const uint64_t m = wave_ballot(is_meas);
    const uint32_t wv = threadIdx.x / kWave;
    if (lane() == 0u) s_w[wv] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    uint32_t before = 0u;
    for (uint32_t w = 0; w < wv; ++w) before += s_w[w];
    if (is_meas) {
        const uint32_t pos = block_off[blockIdx.x] + before + prefix_count(m);
        if (pos < cap) {
#pragma unroll
            for (int q = 0; q < kFeatures; ++q) {
                features[static_cast<size_t>(pos) * kFeatures + q] = P.log_feat[static_cast<size_t>(g) * kFeatures + q];
            }
            labels[pos] = P.log_label[g];
        }
    }
}

hipError_t launch_label_best(const ckf_params& P, const launch_shape& s, hipStream_t stream) {
    hipLaunchKernelGGL(k_label_best, dim3(s.build_blocks), dim3(kBlock), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_log_count(const ckf_params& P, uint32_t n, uint32_t* block_n, hipStream_t stream) {
    hipLaunchKernelGGL(k_log_count, dim3((n + 1023u) / 1024u), dim3(1024), 0, stream, P, n, block_n);
    return hipGetLastError();
}

hipError_t launch_log_write(const ckf_params& P, uint32_t n, const uint32_t* block_off,
                            float* features, uint8_t* labels, uint32_t cap, hipStream_t stream) {
    hipLaunchKernelGGL(k_log_write, dim3((n + 1023u) / 1024u), dim3(1024), 0, stream, P, n,
                       block_off, features, labels, cap);
    return hipGetLastError();
}

hipError_t launch_dedup(const ckf_params& P, const launch_shape& s, hipStream_t stream) {
    hipLaunchKernelGGL(k_dedup_insert, dim3(s.dedup_blocks), dim3(kBlock), 0, stream, P);
    hipLaunchKernelGGL(k_dedup_mark, dim3(s.dedup_blocks), dim3(kBlock), 0, stream, P);
    hipLaunchKernelGGL(k_dedup_emit, dim3(s.emit_blocks), dim3(kBlock), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_propagate(const ckf_params& P, const launch_shape& s, hipStream_t stream,
                            bool fused_advance) {
    const hipError_t e = launch_propagate_work(P, s, stream);
    if (e != hipSuccess) {
        return e;
    }
    return launch_prop_emit(P, s, stream, fused_advance);
}

hipError_t launch_prop_emit(const ckf_params& P, const launch_shape& s, hipStream_t stream,
                            bool fused_advance) {
    hipLaunchKernelGGL(k_prop_emit, dim3(s.emit_blocks), dim3(kBlock), 0, stream, P,
                       fused_advance ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_propagate_work(const ckf_params& P, const launch_shape& s, hipStream_t stream) {
    switch (s.prop_team) {
        case 16:
            hipLaunchKernelGGL(k_propagate<16>, dim3(s.prop_blocks), dim3(kBlock), 0, stream, P);
            break;
        case 8:
            hipLaunchKernelGGL(k_propagate<8>, dim3(s.prop_blocks), dim3(kBlock), 0, stream, P);
            break;
        case 4:
            hipLaunchKernelGGL(k_propagate<4>, dim3(s.prop_blocks), dim3(kBlock), 0, stream, P);
            break;
        case 2:
            hipLaunchKernelGGL(k_propagate<2>, dim3(s.prop_blocks), dim3(kBlock), 0, stream, P);
            break;
        default:
            hipLaunchKernelGGL(k_propagate<1>, dim3(s.prop_blocks), dim3(kBlock), 0, stream, P);
            hipLaunchKernelGGL(k_propagate_resume, dim3(s.resume_blocks), dim3(kBlock), 0, stream, P);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_advance(const ckf_params& P, hipStream_t stream) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_ckf_step(const ckf_params& P, const launch_shape& s, hipStream_t stream,
                           bool with_dedup) {
    hipError_t e = launch_find(P, s, stream);
    if (e == hipSuccess && with_dedup) e = launch_dedup(P, s, stream);
    if (e == hipSuccess) e = launch_propagate(P, s, stream, true);
    return e;
}

hipError_t launch_build_count(const ckf_params& P, const launch_shape& s, hipStream_t stream) {
    hipLaunchKernelGGL(k_build_scan, dim3(s.build_blocks), dim3(kBlock), 0, stream, P);
    hipLaunchKernelGGL(k_build_table, dim3(1), dim3(1024), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_build_write(const ckf_params& P, const launch_shape& s, uint32_t n_tips,
                              hipStream_t stream) {
    // pointer chasing back from every tip: as many lanes in flight as there
    // are tips, so the walks overlap instead of queueing in a strided loop
    (void)s;
    const uint32_t want = (n_tips + kBlock - 1u) / kBlock;
    const uint32_t blocks = want < 1u ? 1u : (want > (1u << 20) ? (1u << 20) : want);
    hipLaunchKernelGGL(k_build_write, dim3(blocks), dim3(kBlock), 0, stream, P);
    return hipGetLastError();
}

}  // namespace tamd
