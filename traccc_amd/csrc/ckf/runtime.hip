// Host runtime of the CKF: the C ABI of include/traccc_amd/ckf.h.
//
// What it replaces: the orchestration of traccc's device CKF
// (device/common/src/finding/combinatorial_kalman_filter_algorithm.cpp:52-704),
// which reads sizes back every step.  Here the step loop runs from a device
// control block; the host only checks it once per chunk of steps (and replays
// the chunk as a captured HIP graph), then once more to size the output.
// Link / tip / state buffers that turn out too small are regrown and the call
// is re-run (traccc regrows its link buffer, :310-357).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <map>
#include <mutex>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../common/errors.hpp"
#include "handles.hpp"
#include "kernels.hpp"
#include "traccc_amd/ckf.h"
#include "traccc_amd/ckf_debug.h"
#include "traccc_amd/scorer.h"

namespace tamd {
block_cache& result_cache() {
    static block_cache* c = new block_cache();  // lives until the process ends
    return *c;
}
}  // namespace tamd

namespace {

using tamd::ckf_error;
using tamd::guarded;
using tamd::hip_check;
using tamd::dbuf;
using tamd::dev_copy;
using tamd::result_cache;

void validate_config(const ckf_finding_config& c) {
    // traccc: core/src/finding/combinatorial_kalman_filter_algorithm.cpp:27-31
    if (c.min_track_candidates_per_track == 0u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT,
                        "min_track_candidates_per_track must be at least 1");
    }
    // traccc device ctor: device/common/src/finding/combinatorial_kalman_filter_algorithm.cpp:33-46
    if (!(c.min_step_length_for_next_surface > std::fabs(c.propagation.overstep_tolerance))) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT,
                        "min_step_length_for_next_surface must exceed |overstep_tolerance|");
    }
    if (c.run_smoother == CKF_SMOOTHER_KALMAN) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT,
                        "a Kalman smoother needs a fitter, which this algorithm does not own");
    }
    if (c.max_num_tracks_per_measurement != 0u &&
        !(c.min_measurement_voting_fraction >= 0.f && c.min_measurement_voting_fraction <= 1.f)) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT,
                        "min_measurement_voting_fraction must lie in [0, 1]");
    }
    if (c.max_num_branches_per_surface == 0u || c.max_track_candidates_per_track == 0u ||
        c.max_track_candidates_per_track > 65535u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "branch / candidate limits out of range");
    }
    if (c.propagation.search_window[0] > 1u || c.propagation.search_window[1] > 1u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "search_window is limited to 1 bin per side");
    }
    if (c.max_step_counts_for_next_surface == 0u || !(std::fabs(c.ptc_charge) > 0.f)) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "invalid propagation settings");
    }
}

}  // namespace

struct ckf_algorithm {
    ckf_finding_config cfg{};
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int n_cu = 256;
    uint32_t epoch_next = 1u;
    ckf_run_stats stats{};
    // capacities
    uint32_t in_cap = 0, state_cap = 0, flt_cap = 0, link_cap = 0, tip_cap = 0, lb_cap = 0;
    uint32_t seed_cap = 0, meas_cap = 0, seg_cap = 0, ev_cap = 0;
    // work buffers
    dbuf state[2], flt, stage_in, stage_slot, links, live, tips, seed_first, seed_base;
    dbuf alive_prefix, dedup_head, dedup_next, lb_lo[tamd::kLookbacks], lb_hi[tamd::kLookbacks], ctl;
    dbuf tip_keep, tip_rank, tip_cprefix, seg_count, seg_ccount, seg_base, seg_cbase, seg_off,
        seg_coff, ev_toff, ev_coff;
    // measurement votes (max_num_tracks_per_measurement > 0)
    dbuf vote_slots, vote_key, vote_keep, vote_rank, vote_cpos, vote_misc;
    // scorer training data (branch log) and scorer-pruned finding
    dbuf feat_slot, log_feat, log_label, seed_best, list_m, list_c, list_n, list_row,
        find_count, seed_run_min, list_seed, prop_flag;
    // propagations deferred to k_propagate_resume (one lane per link)
    dbuf defer;
    uint32_t defer_cap = 0;
    bool logging = false;
    // candidate-verdict log of scored calls (diagnostics)
    dbuf verdict_log, verdict_n;
    uint32_t verdict_cap = 0;
    const ckf_scorer* scorer = nullptr;
    float score_threshold = 0.f;
    uint32_t list_len = 0;
    // progressive filter (run_pkf)
    dbuf pkf_states, pkf_n, pkf_chi2, pkf_ndf, pkf_holes, pkf_cpos, pkf_lb_lo;
    tamd::step_ctl* h_ctl = nullptr;  // pinned mirror
    // captured step chunks: graph[1] with duplicate removal, graph[0] without
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    std::vector<char> graph_key;
    // per-kernel event timing (diagnostics; the step loop then runs uncaptured)
    uint32_t last_steps = 0;  // steps the previous call ran (sizes the first launch)
    uint32_t prop_team = 1;   // lanes per link in k_propagate (chosen per call)
    uint32_t schedule_override = 0;   // fixed lanes per link, 0 = autotuned
    std::map<uint32_t, uint32_t> tuned;  // log2(seeds) -> lanes per link
    bool timing = false;
    double t_find = 0.0, t_dedup = 0.0, t_prop = 0.0, t_other = 0.0, t_find_work = 0.0,
           t_find_emit = 0.0, t_prop_work = 0.0, t_prop_emit = 0.0;
    uint32_t n_prop_launches = 0;

    ~ckf_algorithm() {
        for (auto& g : graph) {
            if (g) (void)hipGraphExecDestroy(g);
        }
        dbuf* all[] = {&state[0], &state[1], &flt, &stage_in, &stage_slot, &links, &live, &tips,
                       &seed_first, &seed_base, &alive_prefix, &dedup_head, &dedup_next, &ctl,
                       &tip_keep, &tip_rank, &tip_cprefix, &seg_count, &seg_ccount, &seg_base,
                       &seg_cbase, &seg_off, &seg_coff, &ev_toff, &ev_coff, &vote_slots,
                       &vote_key, &vote_keep, &vote_rank, &vote_cpos, &vote_misc,
                       &pkf_states, &pkf_n, &pkf_chi2, &pkf_ndf, &pkf_holes, &pkf_cpos,
                       &pkf_lb_lo, &feat_slot, &log_feat,
                       &log_label, &seed_best, &list_m, &list_c, &list_n, &list_row, &find_count,
                       &seed_run_min, &list_seed, &prop_flag, &verdict_log, &verdict_n, &defer};
        for (dbuf* b : all) b->release();
        for (int k = 0; k < tamd::kLookbacks; ++k) {
            lb_lo[k].release();
            lb_hi[k].release();
        }
        if (h_ctl) (void)hipHostFree(h_ctl);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

namespace {

constexpr uint32_t kStepsPerChunk = 4u;

uint32_t parts(uint32_t n) { return (n + 63u) / 64u; }

void ensure_buffers(ckf_algorithm& A, const ckf_device_event& ev) {
    const uint32_t N = A.cfg.max_num_branches_per_surface;
    A.flt_cap = static_cast<uint32_t>(std::min<uint64_t>(uint64_t(A.in_cap) * N, 0x3fffffffu));
    A.state_cap = A.flt_cap;
    // look-back partitions: k_find_emit takes 64 inputs per wave,
    // k_propagate 64 / team links per wave (teams of up to 16 lanes), the
    // other kernels 64 items per wave (the 16-input bound is kept as slack)
    A.lb_cap = std::max({(A.in_cap + 15u) / 16u, (A.state_cap + 3u) / 4u, parts(A.tip_cap)}) + 1u;
    const size_t st = size_t(tamd::kStateCols) * 4u;
    A.state[0].ensure(st * A.state_cap);
    A.state[1].ensure(st * A.state_cap);
    A.flt.ensure(st * A.flt_cap);
    A.stage_in.ensure(sizeof(tamd::stage_head) * size_t(A.in_cap));
    A.stage_slot.ensure(sizeof(tamd::stage_cand) * size_t(A.flt_cap));
    A.links.ensure(sizeof(tamd::dlink) * size_t(A.link_cap));
    A.live.ensure(4u * size_t(A.in_cap));
    A.find_count.ensure(4u * size_t(A.in_cap));
    A.prop_flag.ensure(size_t(A.state_cap));
    // a propagation that needs a second layer search (a few % of them)
    // continues in its own kernel; beyond the buffer it continues in place
    A.defer_cap = std::max<uint32_t>(A.state_cap / 8u, 1024u);
    A.defer.ensure(4u * size_t(tamd::kDeferWords) * A.defer_cap);
    A.tips.ensure(4u * size_t(A.tip_cap));
    A.seed_first.ensure(4u * size_t(std::max(ev.n_seeds, 1u)));
    A.seed_base.ensure(4u * size_t(std::max(ev.n_seeds, 1u)));
    A.alive_prefix.ensure(4u * size_t(A.state_cap));
    {
        const size_t old = A.dedup_head.bytes;
        A.dedup_head.ensure(8u * (size_t(ev.n_meas) + 1u));
        if (A.dedup_head.bytes != old) {
            // epoch 0 tags no step: fresh heads read as empty lists
            hip_check(hipMemsetAsync(A.dedup_head.p, 0, A.dedup_head.bytes, A.stream), "memset");
        }
    }
    A.dedup_next.ensure(4u * size_t(A.state_cap));
    for (int k = 0; k < tamd::kLookbacks; ++k) {
        for (dbuf* w : {&A.lb_lo[k], &A.lb_hi[k]}) {
            const size_t old = w->bytes;
            w->ensure(8u * size_t(A.lb_cap));
            if (w->bytes != old) {
                // fresh descriptors carry epoch 0, which no step uses
                hip_check(hipMemsetAsync(w->p, 0, w->bytes, A.stream), "memset");
            }
        }
    }
    A.ctl.ensure(sizeof(tamd::step_ctl));
    A.tip_keep.ensure(4u * size_t(A.tip_cap));
    A.tip_rank.ensure(4u * size_t(A.tip_cap));
    A.tip_cprefix.ensure(4u * size_t(A.tip_cap));
    const size_t seg = size_t(ev.n_events) * A.cfg.max_track_candidates_per_track;
    if (seg >= (size_t(1) << 31)) {
        // build_tracks keys (event, step) segments by a 32-bit index
        throw ckf_error(CKF_ERR_UNSUPPORTED, "n_events * max_track_candidates_per_track >= 2^31");
    }
    A.seg_count.ensure(4u * seg);
    A.seg_ccount.ensure(4u * seg);
    A.seg_base.ensure(4u * seg);
    A.seg_cbase.ensure(4u * seg);
    A.seg_off.ensure(4u * seg);
    A.seg_coff.ensure(4u * seg);
    A.ev_toff.ensure(4u * (size_t(ev.n_events) + 1u));
    A.ev_coff.ensure(4u * (size_t(ev.n_events) + 1u));
    if (A.logging) {
        A.feat_slot.ensure(4u * 12u * size_t(A.flt_cap));
        A.log_feat.ensure(4u * 12u * size_t(A.link_cap));
        A.log_label.ensure(size_t(A.link_cap));
        A.seed_best.ensure(8u * size_t(std::max(ev.n_seeds, 1u)));
    }
    if (A.logging || A.scorer != nullptr) {
        A.seed_run_min.ensure(8u * size_t(std::max(ev.n_seeds, 1u)));
        // every key of a real step compares below all-ones
        hip_check(hipMemsetAsync(A.seed_run_min.p, 0xff, A.seed_run_min.bytes, A.stream), "memset");
    }
    if (A.scorer != nullptr) {
        // the N + 1 best candidates by (chi2, index): the plain CKF's own
        // selection plus one in place of a failed gain update (the plain
        // find walks on until N updates succeed), so keeping every candidate
        // is the plain CKF; refilling pruned places from longer lists
        // (N + 2, N + 4) measured the same C5 efficiency and 1-4 % slower
        // calls
        A.list_len = std::min(A.cfg.max_num_branches_per_surface + 1u, 16u);
        const size_t rows = size_t(A.in_cap) * A.list_len;
        A.list_seed.ensure(4u * size_t(A.in_cap));
        A.list_m.ensure(4u * rows);
        A.list_c.ensure(4u * rows);
        A.list_row.ensure(2u * 16u * rows);
        A.list_n.ensure(4u * size_t(A.in_cap));
    }
    if (A.h_ctl == nullptr) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&A.h_ctl), sizeof(tamd::step_ctl),
                                hipHostMallocDefault),
                  "hipHostMalloc");
    }
}

tamd::ckf_params make_params(ckf_algorithm& A, const ckf_device_detector& D,
                             const ckf_device_event& ev) {
    tamd::ckf_params P;
    std::memset(&P, 0, sizeof(P));  // padding too: P's bytes key the captured graph
    const ckf_finding_config& c = A.cfg;
    P.sfs = D.d_sfs;
    P.layers = D.d_layers;
    P.n_layers = static_cast<uint32_t>(D.h_layers.size());
    P.groups = D.d_groups;
    P.n_groups = static_cast<uint32_t>(D.h_groups.size());
    P.defer = A.defer.as<uint32_t>();
    P.defer_cap = A.defer_cap;
    P.n_surfaces = static_cast<uint32_t>(D.h_sfs.size());
    P.field = tamd::make_field_view<float>(D.field, D.d_map);
    P.nav = tamd::make_nav_config<float>(c);
    P.n_branch_surface = c.max_num_branches_per_surface;
    P.n_branch_seed = c.max_num_branches_per_seed;
    P.min_cands = c.min_track_candidates_per_track;
    P.max_cands = c.max_track_candidates_per_track;
    P.max_skip = c.max_num_skipping_per_cand;
    P.max_consec = c.max_num_consecutive_skipped;
    P.dup_min_len = c.duplicate_removal_minimum_length;
    P.chi2_max = c.chi2_max;
    P.n_events = ev.n_events;
    P.n_seeds = ev.n_seeds;
    P.n_meas = ev.n_meas;
    P.meas_pos = ev.d_pos;
    P.meas_var = ev.d_var;
    P.meas_info = ev.d_info;
    P.meas_ranges = ev.d_ranges;
    P.seed_event = ev.d_seed_event;
    P.seed_surface = ev.d_seed_surface;
    P.seed_vec = ev.d_seed_vec;
    P.seed_cov = ev.d_seed_cov;
    P.state[0] = A.state[0].as<float>();
    P.state[1] = A.state[1].as<float>();
    P.state_cap = A.state_cap;
    P.flt = A.flt.as<float>();
    P.stage_in = A.stage_in.as<tamd::stage_head>();
    P.stage_slot = A.stage_slot.as<tamd::stage_cand>();
    P.flt_cap = A.flt_cap;
    P.in_cap = A.in_cap;
    P.links = A.links.as<tamd::dlink>();
    P.link_cap = A.link_cap;
    P.live = A.live.as<uint32_t>();
    P.find_count = A.find_count.as<uint32_t>();
    P.prop_flag = A.prop_flag.as<uint8_t>();
    P.tips = A.tips.as<uint32_t>();
    P.tip_cap = A.tip_cap;
    P.seed_first = A.seed_first.as<uint32_t>();
    P.seed_base = A.seed_base.as<uint32_t>();
    P.alive_prefix = A.alive_prefix.as<uint32_t>();
    P.dedup_head = A.dedup_head.as<uint64_t>();
    P.dedup_next = A.dedup_next.as<uint32_t>();
    for (int k = 0; k < tamd::kLookbacks; ++k) {
        P.lb_lo[k] = A.lb_lo[k].as<uint64_t>();
        P.lb_hi[k] = A.lb_hi[k].as<uint64_t>();
    }
    P.lb_cap = A.lb_cap;
    P.ctl = A.ctl.as<tamd::step_ctl>();
    P.tip_keep = A.tip_keep.as<uint32_t>();
    P.tip_rank = A.tip_rank.as<uint32_t>();
    P.tip_cprefix = A.tip_cprefix.as<uint32_t>();
    P.seg_count = A.seg_count.as<uint32_t>();
    P.seg_ccount = A.seg_ccount.as<uint32_t>();
    P.seg_base = A.seg_base.as<uint32_t>();
    P.seg_cbase = A.seg_cbase.as<uint32_t>();
    P.seg_off = A.seg_off.as<uint32_t>();
    P.seg_coff = A.seg_coff.as<uint32_t>();
    P.event_track_off = A.ev_toff.as<uint32_t>();
    P.event_const_off = A.ev_coff.as<uint32_t>();
    P.event_meas_off = ev.d_meas_off;
    P.event_win_var = ev.d_win_var;
    if (A.logging) {
        P.feat_slot = A.feat_slot.as<float>();
        P.log_feat = A.log_feat.as<float>();
        P.log_label = A.log_label.as<uint8_t>();
        P.seed_best = A.seed_best.as<uint64_t>();
    }
    if (A.logging || A.scorer != nullptr) {
        P.seed_run_min = A.seed_run_min.as<uint64_t>();
    }
    if (A.scorer != nullptr) {
        P.scored = 1u;
        P.list_seed = A.list_seed.as<uint32_t>();
        P.list_len = A.list_len;
        P.list_m = A.list_m.as<uint32_t>();
        P.list_c = A.list_c.as<float>();
        P.list_n = A.list_n.as<uint32_t>();
        P.list_row = A.list_row.as<short>();
        P.scorer = tamd::scorer_view(A.scorer);
        P.score_threshold = A.score_threshold;
        if (A.verdict_cap > 0u) {
            A.verdict_log.ensure(8u * size_t(A.verdict_cap));
            A.verdict_n.ensure(4u);
            P.verdict_log = A.verdict_log.as<uint64_t>();
            P.verdict_n = A.verdict_n.as<uint32_t>();
            P.verdict_cap = A.verdict_cap;
        }
    }
    return P;
}

tamd::launch_shape make_shape(const ckf_algorithm& A) {
    const uint32_t max_blocks = static_cast<uint32_t>(A.n_cu) * 8u;
    auto blocks = [&](uint32_t cap) {
        const uint32_t b = (parts(cap) + 3u) / 4u;
        return std::max(1u, std::min(b, max_blocks));
    };
    tamd::launch_shape s;
    // grids are upper bounds for the largest step: a wave beyond a step's
    // partitions leaves at once.  The find is sized for 16 inputs per wave
    // (its two-lane teams hold 32, the scored update 64).
    s.find_blocks = std::max(1u, std::min((((A.in_cap + 15u) / 16u) + 3u) / 4u, max_blocks));
    s.dedup_blocks = blocks(A.state_cap);
    s.prop_team = A.prop_team;
    const uint32_t prop_parts = (A.state_cap + 64u / s.prop_team - 1u) / (64u / s.prop_team);
    s.prop_blocks = std::max(1u, std::min((prop_parts + 3u) / 4u, max_blocks));
    s.resume_blocks = std::max(1u, std::min((parts(A.defer_cap) + 3u) / 4u, max_blocks));
    // scorer-pruned finding: 16 inputs per wave listing, 64 per scoring wave
    s.list_blocks = std::max(1u, std::min((((A.in_cap + 15u) / 16u) + 3u) / 4u, max_blocks));
    s.score_blocks = std::max(1u, std::min((((A.in_cap + 63u) / 64u) + 3u) / 4u, max_blocks));
    s.build_blocks = blocks(A.tip_cap);
    // compaction waves do little per partition: four blocks per CU
    s.emit_blocks = std::min(blocks(A.in_cap), static_cast<uint32_t>(A.n_cu) * 4u);
    return s;
}

// Step loop with HIP events around every kernel group, one step at a time.
uint32_t run_steps_timed(ckf_algorithm& A, const tamd::ckf_params& P, const tamd::launch_shape& s) {
    hipEvent_t ev[7];
    for (auto& e : ev) hip_check(hipEventCreate(&e), "event");
    A.t_find = A.t_dedup = A.t_prop = A.t_other = A.t_find_work = A.t_find_emit = 0.0;
    A.t_prop_work = A.t_prop_emit = 0.0;
    A.n_prop_launches = 0u;
    for (uint32_t k = 0; k < A.cfg.max_track_candidates_per_track + 1u; ++k) {
        hip_check(hipEventRecord(ev[0], A.stream), "record");
        hip_check(tamd::launch_find_work(P, s, A.stream), "find");
        hip_check(hipEventRecord(ev[5], A.stream), "record");
        hip_check(tamd::launch_find_emit(P, s, A.stream), "find emit");
        hip_check(hipEventRecord(ev[1], A.stream), "record");
        hip_check(tamd::launch_dedup(P, s, A.stream), "dedup");
        hip_check(hipEventRecord(ev[2], A.stream), "record");
        hip_check(tamd::launch_propagate_work(P, s, A.stream), "propagate");
        hip_check(hipEventRecord(ev[6], A.stream), "record");
        hip_check(tamd::launch_prop_emit(P, s, A.stream, false), "propagate emit");
        hip_check(hipEventRecord(ev[3], A.stream), "record");
        hip_check(tamd::launch_advance(P, A.stream), "advance");
        hip_check(hipEventRecord(ev[4], A.stream), "record");
        hip_check(hipMemcpyAsync(A.h_ctl, P.ctl, sizeof(tamd::step_ctl), hipMemcpyDeviceToHost,
                                 A.stream),
                  "ctl readback");
        hip_check(hipStreamSynchronize(A.stream), "sync");
        float ms[4], ms_work = 0.f, ms_emit = 0.f;
        for (int j = 0; j < 4; ++j) hip_check(hipEventElapsedTime(&ms[j], ev[j], ev[j + 1]), "elapsed");
        hip_check(hipEventElapsedTime(&ms_work, ev[0], ev[5]), "elapsed");
        hip_check(hipEventElapsedTime(&ms_emit, ev[5], ev[1]), "elapsed");
        A.t_find_work += ms_work;
        A.t_find_emit += ms_emit;
        float ms_pw = 0.f, ms_pe = 0.f;
        hip_check(hipEventElapsedTime(&ms_pw, ev[2], ev[6]), "elapsed");
        hip_check(hipEventElapsedTime(&ms_pe, ev[6], ev[3]), "elapsed");
        A.t_prop_work += ms_pw;
        A.t_prop_emit += ms_pe;
        A.t_find += ms[0];
        A.t_dedup += ms[1];
        A.t_prop += ms[2];
        A.t_other += ms[3];
        A.n_prop_launches += 1u;
        if (A.h_ctl->done) break;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    A.epoch_next = A.h_ctl->epoch + 2u;
    return A.h_ctl->overflow;
}

// Runs the step loop once; returns the overflow bits (0 = complete).
uint32_t run_steps(ckf_algorithm& A, const tamd::ckf_params& P, const tamd::launch_shape& s,
                   const ckf_device_event& ev) {
    tamd::step_ctl init{};
    init.epoch = A.epoch_next;
    init.n_in = ev.n_seeds;
    init.done = (ev.n_seeds == 0u) ? 1u : 0u;
    std::memcpy(A.h_ctl, &init, sizeof(init));
    hip_check(hipMemcpyAsync(P.ctl, A.h_ctl, sizeof(init), hipMemcpyHostToDevice, A.stream),
              "ctl upload");
    if (P.verdict_n != nullptr) {
        hip_check(hipMemsetAsync(P.verdict_n, 0, 4u, A.stream), "memset");
    }
    hip_check(hipStreamSynchronize(A.stream), "sync");

    if (A.timing) {
        return run_steps_timed(A, P, s);
    }
    // capture one chunk of steps (with and without duplicate removal);
    // re-capture when the launch parameters change
    std::vector<char> key(sizeof(P) + sizeof(s));
    std::memcpy(key.data(), &P, sizeof(P));
    std::memcpy(key.data() + sizeof(P), &s, sizeof(s));
    if (A.graph[0] == nullptr || key != A.graph_key) {
        for (int w = 0; w < 2; ++w) {
            if (A.graph[w]) {
                (void)hipGraphExecDestroy(A.graph[w]);
                A.graph[w] = nullptr;
            }
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(A.stream, hipStreamCaptureModeThreadLocal),
                      "begin capture");
            for (uint32_t k = 0; k < kStepsPerChunk; ++k) {
                hip_check(tamd::launch_ckf_step(P, s, A.stream, w == 1), "step launch");
            }
            hip_check(hipStreamEndCapture(A.stream, &g), "end capture");
            hip_check(hipGraphInstantiate(&A.graph[w], g, nullptr, nullptr, 0), "graph instantiate");
            (void)hipGraphDestroy(g);
        }
        A.graph_key = key;
    }
    uint32_t launched = 0u;
    // launch as many chunks as the previous call needed before the first check
    uint32_t chunks = std::max(1u, (A.last_steps + kStepsPerChunk - 1u) / kStepsPerChunk);
    for (;;) {
        for (uint32_t k = 0; k < chunks; ++k) {
            // a chunk whose last step is below duplicate_removal_minimum_length
            // runs without the duplicate-removal kernels
            const uint32_t last_step = launched + (k + 1u) * kStepsPerChunk - 1u;
            const int w = last_step >= A.cfg.duplicate_removal_minimum_length ? 1 : 0;
            hip_check(hipGraphLaunch(A.graph[w], A.stream), "graph launch");
        }
        launched += chunks * kStepsPerChunk;
        chunks = 1u;
        hip_check(hipMemcpyAsync(A.h_ctl, P.ctl, sizeof(tamd::step_ctl), hipMemcpyDeviceToHost,
                                 A.stream),
                  "ctl readback");
        hip_check(hipStreamSynchronize(A.stream), "step sync");
        if (A.h_ctl->done || launched >= A.cfg.max_track_candidates_per_track + kStepsPerChunk) {
            break;
        }
    }
    A.epoch_next = A.h_ctl->epoch + 2u;
    A.last_steps = A.h_ctl->n_steps_run;
    return A.h_ctl->overflow;
}

// Grow the buffers that overflowed: at least 4x, and at least twice what the
// aborted step already needed (the control block holds the true counts).
void grow(ckf_algorithm& A, uint32_t overflow) {
    const tamd::step_ctl& c = *A.h_ctl;
    auto up = [](uint32_t cur, uint64_t need) {
        const uint64_t v = std::max<uint64_t>(uint64_t(cur) * 4u, need * 2u);
        return static_cast<uint32_t>(std::min<uint64_t>(v, 0x3fffffffu));
    };
    const uint64_t N = A.cfg.max_num_branches_per_surface;
    if (overflow & (2u | 4u)) A.in_cap = up(A.in_cap, c.n_in);
    if (overflow & 16u) A.in_cap = up(A.in_cap, std::max(c.n_in, c.n_live));
    if (overflow & 1u) {
        A.link_cap = up(A.link_cap, uint64_t(c.link_base) + c.n_links + uint64_t(c.n_in) * N);
        A.tip_cap = std::max(A.tip_cap, A.link_cap / 2u);
    }
    if (overflow & 8u) A.tip_cap = up(A.tip_cap, uint64_t(c.n_tips) + c.n_tips_step);
}

// Measurement-vote pruning of the built tracks in T: the kept tracks move to
// fresh result blocks, which T takes; the built blocks move to `old`, which
// the caller keeps until the stream has passed the vote kernels.  The new
// per-event offsets are left in A.vote_misc after the two totals.
void run_votes(ckf_algorithm& A, const tamd::ckf_params& P, const ckf_device_event& ev,
               ckf_device_tracks& T, ckf_device_tracks& K2) {
    const ckf_finding_config& c = A.cfg;
    const uint32_t nt = T.n_tracks;
    const uint32_t nc = T.n_const;
    const uint32_t K = c.max_num_tracks_per_measurement;
    const size_t slot_bytes = 8u * size_t(K) * std::max<size_t>(ev.n_meas, 1u);
    A.vote_slots.ensure(slot_bytes);
    A.vote_keep.ensure(4u * size_t(nt));
    A.vote_rank.ensure(4u * size_t(nt));
    A.vote_cpos.ensure(4u * size_t(nt));
    A.vote_misc.ensure(4u * (2u + 2u * (size_t(ev.n_events) + 1u)));
    hip_check(hipMemsetAsync(A.vote_slots.p, 0, slot_bytes, A.stream), "memset");
    hip_check(hipMemsetAsync(&P.ctl->ticket[1].v, 0, 4u, A.stream), "memset");

    const size_t n1 = std::max<size_t>(nt, 1);
    const size_t sizes[10] = {2u * n1, 4u * n1, 24u * n1, 144u * n1, 4u * n1, 4u * n1,
                              4u * n1, 4u * n1, 4u * (size_t(nt) + 1u),
                              4u * std::max<size_t>(nc, 1)};
    void** slots[10] = {reinterpret_cast<void**>(&K2.outcome), reinterpret_cast<void**>(&K2.surface),
                        reinterpret_cast<void**>(&K2.vec), reinterpret_cast<void**>(&K2.cov),
                        reinterpret_cast<void**>(&K2.ndf), reinterpret_cast<void**>(&K2.chi2),
                        reinterpret_cast<void**>(&K2.pval), reinterpret_cast<void**>(&K2.nholes),
                        reinterpret_cast<void**>(&K2.offsets), reinterpret_cast<void**>(&K2.cons)};
    for (int i = 0; i < 10; ++i) {
        *slots[i] = result_cache().get(sizes[i]);
        K2.track_bytes[i] = sizes[i];
    }

    tamd::vote_params V{};
    V.ctl = P.ctl;
    V.n_tracks = nt;
    V.n_events = ev.n_events;
    V.K = K;
    V.min_fraction = c.min_measurement_voting_fraction;
    V.track_off = P.event_track_off;
    V.event_meas_off = P.event_meas_off;
    V.offsets = T.offsets;
    V.cons = T.cons;
    V.key = A.vote_key.as<uint64_t>();
    V.slots = A.vote_slots.as<uint64_t>();
    V.keep = A.vote_keep.as<uint32_t>();
    V.rank = A.vote_rank.as<uint32_t>();
    V.cpos = A.vote_cpos.as<uint32_t>();
    V.totals = A.vote_misc.as<uint32_t>();
    V.new_track_off = V.totals + 2u;
    V.new_const_off = V.new_track_off + (ev.n_events + 1u);
    V.lb_lo = A.lb_lo[1].as<uint64_t>();
    V.lb_hi = A.lb_hi[1].as<uint64_t>();
    V.outcome = T.outcome;
    V.surface = T.surface;
    V.vec = T.vec;
    V.cov = T.cov;
    V.ndf = T.ndf;
    V.chi2 = T.chi2;
    V.pval = T.pval;
    V.nholes = T.nholes;
    V.o_outcome = K2.outcome;
    V.o_surface = K2.surface;
    V.o_vec = K2.vec;
    V.o_cov = K2.cov;
    V.o_ndf = K2.ndf;
    V.o_chi2 = K2.chi2;
    V.o_pval = K2.pval;
    V.o_nholes = K2.nholes;
    V.o_offsets = K2.offsets;
    V.o_cons = K2.cons;
    hip_check(tamd::launch_votes(V, static_cast<uint32_t>(A.n_cu), A.stream), "votes");
    std::swap(T.outcome, K2.outcome);
    std::swap(T.surface, K2.surface);
    std::swap(T.vec, K2.vec);
    std::swap(T.cov, K2.cov);
    std::swap(T.ndf, K2.ndf);
    std::swap(T.chi2, K2.chi2);
    std::swap(T.pval, K2.pval);
    std::swap(T.nholes, K2.nholes);
    std::swap(T.offsets, K2.offsets);
    std::swap(T.cons, K2.cons);
    for (int i = 0; i < 10; ++i) std::swap(T.track_bytes[i], K2.track_bytes[i]);
}

// Propagation schedule autotuner: lanes per link in k_propagate.  Every
// schedule gives the same tracks; which is fastest depends on how many links
// a step holds against the machine's lanes (one lane per link fills the chip
// once tens of thousands of links are live; a small event runs its links'
// navigation on teams of lanes instead).  The first call of each size class
// (log2 of the seed count) times the step loop once per schedule on its own
// input and keeps the fastest; later calls of the class reuse it.  A fixed
// schedule (ckf_debug_set_schedule) bypasses the tuner.
template <typename Loop>
uint32_t choose_schedule(ckf_algorithm& A, uint32_t n_seeds, Loop& step_loop) {
    if (A.schedule_override != 0u) {
        return A.schedule_override;
    }
    uint32_t cls = 0u;
    while ((1u << cls) < n_seeds && cls < 31u) ++cls;
    const auto it = A.tuned.find(cls);
    if (it != A.tuned.end()) {
        return it->second;
    }
    const uint32_t teams[5] = {1u, 2u, 4u, 8u, 16u};
    uint32_t best = 1u;
    double best_s = 1e300;
    for (uint32_t t : teams) {
        A.prop_team = t;
        // the faster of two runs: a schedule's first run also captures its
        // step graph, which would bias the choice against whichever runs first
        for (int rep = 0; rep < 2; ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            step_loop();  // ends with a stream synchronisation
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (dt < best_s) {
                best_s = dt;
                best = t;
            }
        }
    }
    A.tuned[cls] = best;
    return best;
}

// The operator's dispatch (core/src/finding/combinatorial_kalman_filter_algorithm.cpp:47-57):
// the progressive filter runs when asked for and at most one branch is kept.
bool use_pkf(const ckf_finding_config& c) {
    return c.run_pkf != 0u &&
           (c.max_num_branches_per_seed == 1u || c.max_num_branches_per_surface == 1u);
}

ckf_device_tracks* find_tracks_pkf(ckf_algorithm& A, const ckf_device_detector& D,
                                   const ckf_device_event& ev) {
    const ckf_finding_config& c = A.cfg;
    const uint32_t ns = ev.n_seeds;
    const size_t n1 = std::max<size_t>(ns, 1u);
    const uint32_t max_states = c.max_track_candidates_per_track;
    A.pkf_states.ensure(4u * n1 * max_states);
    A.pkf_n.ensure(4u * n1);
    A.pkf_chi2.ensure(4u * n1);
    A.pkf_ndf.ensure(4u * n1);
    A.pkf_holes.ensure(4u * n1);
    A.pkf_cpos.ensure(4u * n1);
    const uint32_t lb = parts(ns) + 1u;
    {
        const size_t old = A.pkf_lb_lo.bytes;
        A.pkf_lb_lo.ensure(8u * size_t(lb));
        if (A.pkf_lb_lo.bytes != old) {
            // fresh descriptors carry epoch 0, which no launch uses
            hip_check(hipMemsetAsync(A.pkf_lb_lo.p, 0, A.pkf_lb_lo.bytes, A.stream), "memset");
        }
    }
    A.ctl.ensure(sizeof(tamd::step_ctl));
    A.ev_toff.ensure(4u * (size_t(ev.n_events) + 1u));
    A.ev_coff.ensure(4u * (size_t(ev.n_events) + 1u));
    if (A.h_ctl == nullptr) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&A.h_ctl), sizeof(tamd::step_ctl),
                                hipHostMallocDefault),
                  "hipHostMalloc");
    }
    A.stats = ckf_run_stats{};
    tamd::ckf_params P = make_params(A, D, ev);
    tamd::pkf_params Q{};
    Q.states = A.pkf_states.as<uint32_t>();
    Q.n_states = A.pkf_n.as<uint32_t>();
    Q.chi2 = A.pkf_chi2.as<float>();
    Q.ndf = A.pkf_ndf.as<uint32_t>();
    Q.holes = A.pkf_holes.as<uint32_t>();
    Q.cpos = A.pkf_cpos.as<uint32_t>();
    Q.max_states = max_states;
    Q.lb_lo = A.pkf_lb_lo.as<uint64_t>();
    Q.lb_hi = nullptr;

    tamd::step_ctl init{};
    init.epoch = A.epoch_next;
    std::memcpy(A.h_ctl, &init, sizeof(init));
    hip_check(hipMemcpyAsync(P.ctl, A.h_ctl, sizeof(init), hipMemcpyHostToDevice, A.stream),
              "ctl upload");
    if (ns > 0u) {
        hip_check(tamd::launch_pkf(P, Q, static_cast<uint32_t>(A.n_cu), A.stream), "pkf");
    }
    hip_check(hipMemcpyAsync(A.h_ctl, P.ctl, sizeof(tamd::step_ctl), hipMemcpyDeviceToHost,
                             A.stream),
              "ctl readback");
    hip_check(hipStreamSynchronize(A.stream), "pkf sync");
    A.epoch_next = A.h_ctl->epoch + 2u;
    const uint32_t nc = ns > 0u ? A.h_ctl->n_constituents : 0u;
    A.stats.n_steps = 1u;
    A.stats.n_tips = ns;

    auto* T = new ckf_device_tracks();
    try {
        T->n_events = ev.n_events;
        T->n_tracks = ns;
        T->n_const = nc;
        T->ids = D.ids;
        const size_t sizes[10] = {2u * n1, 4u * n1, 24u * n1, 144u * n1, 4u * n1, 4u * n1,
                                  4u * n1, 4u * n1, 4u * (size_t(ns) + 1u),
                                  4u * std::max<size_t>(nc, 1)};
        void** slots[10] = {reinterpret_cast<void**>(&T->outcome), reinterpret_cast<void**>(&T->surface),
                            reinterpret_cast<void**>(&T->vec), reinterpret_cast<void**>(&T->cov),
                            reinterpret_cast<void**>(&T->ndf), reinterpret_cast<void**>(&T->chi2),
                            reinterpret_cast<void**>(&T->pval), reinterpret_cast<void**>(&T->nholes),
                            reinterpret_cast<void**>(&T->offsets), reinterpret_cast<void**>(&T->cons)};
        for (int i = 0; i < 10; ++i) {
            *slots[i] = result_cache().get(sizes[i]);
            T->track_bytes[i] = sizes[i];
        }
        P.out_outcome = T->outcome;
        P.out_surface = T->surface;
        P.out_vec = T->vec;
        P.out_cov = T->cov;
        P.out_ndf = T->ndf;
        P.out_chi2 = T->chi2;
        P.out_pval = T->pval;
        P.out_nholes = T->nholes;
        P.out_offsets = T->offsets;
        P.out_const = T->cons;
        P.out_cap = ns;
        P.out_const_cap = nc;
        T->ev_track_off.assign(ev.seed_off.begin(), ev.seed_off.end());
        T->ev_const_off.assign(ev.n_events + 1u, 0u);
        if (ns > 0u) {
            hip_check(tamd::launch_pkf_write(P, Q, static_cast<uint32_t>(A.n_cu), A.stream),
                      "pkf write");
            hip_check(hipMemcpyAsync(T->ev_const_off.data(), P.event_const_off,
                                     4u * ev.n_events, hipMemcpyDeviceToHost, A.stream),
                      "offsets");
            hip_check(hipStreamSynchronize(A.stream), "pkf write sync");
        } else {
            hip_check(hipMemcpyAsync(T->offsets, T->ev_const_off.data(), 4u, hipMemcpyHostToDevice,
                                     A.stream),
                      "offsets");
            hip_check(hipStreamSynchronize(A.stream), "pkf write sync");
        }
        // events without seeds start where the next event starts
        T->ev_const_off[ev.n_events] = nc;
        for (uint32_t e = ev.n_events; e-- > 0u;) {
            if (ev.seed_off[e] == ev.seed_off[e + 1u]) {
                T->ev_const_off[e] = T->ev_const_off[e + 1u];
            }
        }
    } catch (...) {
        delete T;
        throw;
    }
    return T;
}

ckf_device_tracks* find_tracks(ckf_algorithm& A, const ckf_device_detector& D,
                               const ckf_device_event& ev) {
    const ckf_finding_config& c = A.cfg;
    if (use_pkf(c)) {
        return find_tracks_pkf(A, D, ev);
    }
    // initial capacities; grown and re-run on overflow
    const uint32_t ns = std::max(ev.n_seeds, 1u);
    const uint64_t branchy = std::min<uint64_t>(c.max_num_branches_per_seed,
                                                uint64_t(c.max_num_branches_per_surface) *
                                                    c.max_num_branches_per_surface);
    A.in_cap = std::max<uint32_t>(A.in_cap, static_cast<uint32_t>(std::min<uint64_t>(
                                                uint64_t(ns) * std::max<uint64_t>(branchy, 1u),
                                                0x3fffffffu)));
    A.link_cap = std::max<uint32_t>(A.link_cap, static_cast<uint32_t>(std::min<uint64_t>(
                                                    uint64_t(ns) * c.initial_links_per_seed,
                                                    0x3fffffffu)));
    A.tip_cap = std::max<uint32_t>(A.tip_cap, std::max<uint32_t>(ns * 2u, A.link_cap / 4u));
    A.stats = ckf_run_stats{};

    tamd::ckf_params P{};
    tamd::launch_shape shape{};
    auto step_loop = [&]() {
        for (int attempt = 0;; ++attempt) {
            ensure_buffers(A, ev);
            P = make_params(A, D, ev);
            shape = make_shape(A);
            const uint32_t ovf = run_steps(A, P, shape, ev);
            if (ovf == 0u) {
                break;
            }
            if (attempt > 40) {
                throw ckf_error(CKF_ERR_OUT_OF_MEMORY, "CKF buffers kept overflowing");
            }
            grow(A, ovf);
            A.stats.n_regrows += 1u;
        }
    };
    A.prop_team = choose_schedule(A, ns, step_loop);
    step_loop();
    const tamd::step_ctl done = *A.h_ctl;
    A.stats.n_steps = done.n_steps_run;
    A.stats.n_links = done.n_links_total;
    A.stats.n_propagations = done.n_props.v;
    A.stats.n_tips = done.n_tips;
    A.stats.n_scored = done.n_scored;

    // build_tracks: count, size, fill (the build marks logged links on tracks)
    if (A.logging) {
        hip_check(hipMemsetAsync(A.log_label.p, 0, A.log_label.bytes, A.stream), "memset");
        hip_check(hipMemsetAsync(A.seed_best.p, 0, A.seed_best.bytes, A.stream), "memset");
    }
    const size_t seg = size_t(ev.n_events) * c.max_track_candidates_per_track;
    hip_check(hipMemsetAsync(P.seg_count, 0, 4u * seg, A.stream), "memset");
    hip_check(hipMemsetAsync(P.seg_ccount, 0, 4u * seg, A.stream), "memset");
    hip_check(hipMemsetAsync(&P.ctl->ticket[3].v, 0, 4u, A.stream), "memset");
    hip_check(tamd::launch_build_count(P, shape, A.stream), "build count");
    hip_check(hipMemcpyAsync(A.h_ctl, P.ctl, sizeof(tamd::step_ctl), hipMemcpyDeviceToHost,
                             A.stream),
              "ctl readback");
    hip_check(hipStreamSynchronize(A.stream), "build sync");
    A.epoch_next = A.h_ctl->epoch + 2u;

    auto* T = new ckf_device_tracks();
    try {
        const uint32_t nt = A.h_ctl->n_tracks;
        const uint32_t nc = A.h_ctl->n_constituents;
        T->n_events = ev.n_events;
        T->n_tracks = nt;
        T->n_const = nc;
        T->ids = D.ids;
        const size_t n1 = std::max<size_t>(nt, 1);
        const size_t sizes[10] = {2u * n1, 4u * n1, 24u * n1, 144u * n1, 4u * n1, 4u * n1,
                                  4u * n1, 4u * n1, 4u * (size_t(nt) + 1u),
                                  4u * std::max<size_t>(nc, 1)};
        void** slots[10] = {reinterpret_cast<void**>(&T->outcome), reinterpret_cast<void**>(&T->surface),
                            reinterpret_cast<void**>(&T->vec), reinterpret_cast<void**>(&T->cov),
                            reinterpret_cast<void**>(&T->ndf), reinterpret_cast<void**>(&T->chi2),
                            reinterpret_cast<void**>(&T->pval), reinterpret_cast<void**>(&T->nholes),
                            reinterpret_cast<void**>(&T->offsets), reinterpret_cast<void**>(&T->cons)};
        for (int i = 0; i < 10; ++i) {
            *slots[i] = result_cache().get(sizes[i]);
            T->track_bytes[i] = sizes[i];
        }
        P.out_outcome = T->outcome;
        P.out_surface = T->surface;
        P.out_vec = T->vec;
        P.out_cov = T->cov;
        P.out_ndf = T->ndf;
        P.out_chi2 = T->chi2;
        P.out_pval = T->pval;
        P.out_nholes = T->nholes;
        P.out_offsets = T->offsets;
        P.out_const = T->cons;
        P.out_cap = nt;
        P.out_const_cap = nc;
        const bool votes = c.max_num_tracks_per_measurement > 0u && nt > 0u;
        if (votes) {
            A.vote_key.ensure(8u * size_t(nt));
            P.vote_key = A.vote_key.as<uint64_t>();
        }
        if (A.logging) {
            hip_check(tamd::launch_label_best(P, shape, A.stream), "label best");
        }
        hip_check(tamd::launch_build_write(P, shape, A.h_ctl->n_tips, A.stream), "build write");
        const uint32_t* track_off = P.event_track_off;
        const uint32_t* const_off = P.event_const_off;
        ckf_device_tracks built;  // blocks the vote kernels read, released after the sync
        if (votes) {
            run_votes(A, P, ev, *T, built);
            track_off = A.vote_misc.as<uint32_t>() + 2u;
            const_off = track_off + (ev.n_events + 1u);
        }
        T->ev_track_off.resize(ev.n_events + 1u);
        T->ev_const_off.resize(ev.n_events + 1u);
        hip_check(hipMemcpyAsync(T->ev_track_off.data(), track_off, 4u * (ev.n_events + 1u),
                                 hipMemcpyDeviceToHost, A.stream),
                  "offsets");
        hip_check(hipMemcpyAsync(T->ev_const_off.data(), const_off, 4u * (ev.n_events + 1u),
                                 hipMemcpyDeviceToHost, A.stream),
                  "offsets");
        hip_check(hipStreamSynchronize(A.stream), "build write sync");
        if (votes) {
            T->n_tracks = T->ev_track_off[ev.n_events];
            T->n_const = T->ev_const_off[ev.n_events];
        }
    } catch (...) {
        delete T;
        throw;
    }
    return T;
}

ckf_device_event* upload_event(const ckf_device_detector& D, const ckf_measurements* meas,
                               const ckf_bound_params* seeds, uint32_t n_events) {
    if (meas == nullptr || seeds == nullptr || n_events == 0u) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null event input");
    }
    const uint32_t S = static_cast<uint32_t>(D.h_sfs.size());
    auto* E = new ckf_device_event();
    try {
        E->n_events = n_events;
        E->n_surfaces = S;
        E->meas_off.assign(n_events + 1u, 0u);
        E->seed_off.assign(n_events + 1u, 0u);
        for (uint32_t e = 0; e < n_events; ++e) {
            E->meas_off[e + 1] = E->meas_off[e] + meas[e].size;
            E->seed_off[e + 1] = E->seed_off[e] + seeds[e].size;
        }
        const uint32_t nm = E->meas_off[n_events];
        const uint32_t ns = E->seed_off[n_events];
        E->n_meas = nm;
        E->n_seeds = ns;
        std::vector<float2> pos(nm), var(nm);
        std::vector<uint32_t> info(nm), ranges(size_t(n_events) * (S + 1u)), meas_sf(nm, 0u);
    std::vector<float> win_var(n_events, -1.f);
        std::vector<uint32_t> seed_event(ns), seed_surface(ns);
        std::vector<float> seed_vec(size_t(ns) * 6u), seed_cov(size_t(ns) * 36u);
        for (uint32_t e = 0; e < n_events; ++e) {
            const ckf_measurements& M = meas[e];
            const uint32_t off = E->meas_off[e];
            if (M.size > 0u && (M.local_position == nullptr || M.local_variance == nullptr ||
                                M.dimensions == nullptr || M.surface_link == nullptr ||
                                M.subspace == nullptr)) {
                throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "missing measurement column");
            }
            for (uint32_t i = 0; i < M.size; ++i) {
                if (i > 0u && M.surface_link[i] < M.surface_link[i - 1u]) {
                    throw ckf_error(CKF_ERR_PRECONDITION,
                                    "measurements are not grouped by surface");
                }
                const uint32_t d = M.dimensions[i];
                if (d != 1u && d != 2u) {
                    throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "measurement dimension must be 1 or 2");
                }
                const uint32_t s0 = M.subspace[2u * i], s1 = M.subspace[2u * i + 1u];
                if (s0 > 1u || (d == 2u && (s1 > 1u || s1 == s0))) {
                    throw ckf_error(CKF_ERR_UNSUPPORTED,
                                    "measurement subspaces other than loc0 / loc1 are not supported");
                }
                pos[off + i] = make_float2(M.local_position[2u * i], M.local_position[2u * i + 1u]);
                var[off + i] = make_float2(M.local_variance[2u * i], M.local_variance[2u * i + 1u]);
                info[off + i] = d | (s0 << 8) | (s1 << 16);
            }
            uint32_t* row = ranges.data() + size_t(e) * (S + 1u);
            row[0] = off;
            for (uint32_t s = 0; s < S; ++s) {
                const uint64_t id = D.ids[s];
                const uint64_t* b = M.surface_link;
                const uint32_t ub = static_cast<uint32_t>(std::upper_bound(b, b + M.size, id) - b);
                row[s + 1u] = off + ub;
                for (uint32_t m = row[s]; m < row[s + 1u]; ++m) meas_sf[m] = s;
            }
            // every measurement must sit on a detector surface: one whose
            // identifier is unknown would fall into the next surface's range,
            // or (past the last surface) into none
            if (row[S] != off + M.size) {
                throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "measurement on an unknown surface");
            }
            for (uint32_t i = 0; i < M.size; ++i) {
                if (D.ids[meas_sf[off + i]] != M.surface_link[i]) {
                    throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "measurement on an unknown surface");
                }
            }
            // the candidate search may bisect a surface's measurements on
            // loc0 when they are sorted by it (traccc's measurement ordering)
            // and every one of them measures loc0
            {
                bool ok = true;
                float vmax = 0.f;
                for (uint32_t i = 0; i < M.size && ok; ++i) {
                    const float x = M.local_position[2u * i];
                    const float v = M.local_variance[2u * i];
                    ok = M.subspace[2u * i] == 0u && std::isfinite(x) && std::isfinite(v) && v >= 0.f;
                    if (ok && i > 0u && M.surface_link[i] == M.surface_link[i - 1u]) {
                        ok = !(x < M.local_position[2u * (i - 1u)]);
                    }
                    vmax = std::max(vmax, v);
                }
                win_var[e] = ok ? vmax : -1.f;
            }
            const ckf_bound_params& B = seeds[e];
            if (B.size > 0u && (B.surface_link == nullptr || B.vector == nullptr ||
                                B.covariance == nullptr)) {
                throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "missing seed column");
            }
            for (uint32_t i = 0; i < B.size; ++i) {
                const uint32_t g = E->seed_off[e] + i;
                seed_event[g] = e;
                const auto it = std::lower_bound(D.ids.begin(), D.ids.end(), B.surface_link[i]);
                if (it == D.ids.end() || *it != B.surface_link[i]) {
                    throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "seed on an unknown surface");
                }
                seed_surface[g] = static_cast<uint32_t>(it - D.ids.begin());
                std::memcpy(&seed_vec[6u * size_t(g)], B.vector + 6u * size_t(i), 24u);
                std::memcpy(&seed_cov[36u * size_t(g)], B.covariance + 36u * size_t(i), 144u);
            }
        }
        E->d_pos = dev_copy(pos.data(), nm);
        E->d_var = dev_copy(var.data(), nm);
        E->d_info = dev_copy(info.data(), nm);
        E->d_ranges = dev_copy(ranges.data(), ranges.size());
        E->d_seed_event = dev_copy(seed_event.data(), ns);
        E->d_seed_surface = dev_copy(seed_surface.data(), ns);
        E->d_seed_vec = dev_copy(seed_vec.data(), seed_vec.size());
        E->d_seed_cov = dev_copy(seed_cov.data(), seed_cov.size());
        E->d_meas_off = dev_copy(E->meas_off.data(), n_events);
        E->d_meas_sf = dev_copy(meas_sf.data(), nm);
        E->d_win_var = dev_copy(win_var.data(), n_events);
    } catch (...) {
        delete E;
        throw;
    }
    return E;
}

ckf_device_detector* upload_detector(const ckf_detector* det, const ckf_bfield* field) {
    if (det == nullptr || field == nullptr || det->surfaces == nullptr || det->layers == nullptr) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null detector or field");
    }
    if (det->n_layers > 64u) {
        throw ckf_error(CKF_ERR_UNSUPPORTED, "at most 64 navigation layers are supported");
    }
    if (field->kind != CKF_BFIELD_CONSTANT && field->kind != CKF_BFIELD_RZ_MAP) {
        throw ckf_error(CKF_ERR_UNSUPPORTED, "unsupported magnetic field type");
    }
    if (field->kind == CKF_BFIELD_RZ_MAP &&
        (field->map == nullptr || field->n_r < 2u || field->n_z < 2u || !(field->dr > 0.f) ||
         !(field->dz > 0.f))) {
        throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "invalid field map");
    }
    auto* D = new ckf_device_detector();
    try {
        D->h_sfs.assign(det->surfaces, det->surfaces + det->n_surfaces);
        D->h_layers.assign(det->layers, det->layers + det->n_layers);
        D->ids.resize(det->n_surfaces);
        for (uint32_t i = 0; i < det->n_surfaces; ++i) {
            D->ids[i] = det->surfaces[i].identifier;
            if (i > 0u && D->ids[i] <= D->ids[i - 1u]) {
                throw ckf_error(CKF_ERR_INVALID_ARGUMENT,
                                "surface identifiers must increase with the index");
            }
            if (det->surfaces[i].layer >= det->n_layers) {
                throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "surface refers to an unknown layer");
            }
        }
        D->field = *field;
        D->d_sfs = dev_copy(det->surfaces, det->n_surfaces);
        D->d_layers = dev_copy(det->layers, det->n_layers);
        D->h_groups.resize(tamd::kMaxGroups);
        D->h_groups.resize(tamd::build_layer_groups(det->layers, det->n_layers, D->h_groups.data()));
        D->d_groups = dev_copy(D->h_groups.data(), D->h_groups.size());
        if (field->kind == CKF_BFIELD_RZ_MAP) {
            D->d_map = dev_copy(field->map, size_t(2u) * field->n_r * field->n_z);
            D->field.map = D->d_map;
        }
    } catch (...) {
        delete D;
        throw;
    }
    return D;
}

}  // namespace

extern "C" {

uint32_t ckf_abi_version(void) { return CKF_ABI_VERSION; }


ckf_status ckf_algorithm_create(const ckf_finding_config* cfg, ckf_algorithm** out) {
    return guarded([&] {
        if (cfg == nullptr || out == nullptr) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        validate_config(*cfg);
        auto* A = new ckf_algorithm();
        A->cfg = *cfg;
        // a stream of our own (default priority): the legacy null stream
        // cannot be captured into a graph
        if (hipStreamCreateWithFlags(&A->own_stream, hipStreamNonBlocking) != hipSuccess) {
            delete A;
            throw ckf_error(CKF_ERR_DEVICE, "hipStreamCreate failed");
        }
        A->stream = A->own_stream;
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) {
                A->n_cu = prop.multiProcessorCount;
            }
        }
        *out = A;
    });
}

ckf_status ckf_algorithm_destroy(ckf_algorithm* alg) {
    return guarded([&] { delete alg; });
}

ckf_status ckf_algorithm_set_stream(ckf_algorithm* alg, void* hip_stream) {
    return guarded([&] {
        if (alg == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null algorithm");
        alg->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : alg->own_stream;
        for (auto& g : alg->graph) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    });
}

ckf_status ckf_algorithm_last_stats(const ckf_algorithm* alg, ckf_run_stats* stats) {
    return guarded([&] {
        if (alg == nullptr || stats == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null");
        *stats = alg->stats;
    });
}

ckf_status ckf_device_detector_create(const ckf_detector* det, const ckf_bfield* field,
                                      ckf_device_detector** out) {
    return guarded([&] {
        if (out == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null out");
        *out = upload_detector(det, field);
    });
}

ckf_status ckf_device_detector_destroy(ckf_device_detector* det) {
    return guarded([&] { delete det; });
}

ckf_status ckf_device_event_create(const ckf_device_detector* det, const ckf_measurements* meas,
                                   const ckf_bound_params* seeds, uint32_t n_events,
                                   ckf_device_event** out) {
    return guarded([&] {
        if (det == nullptr || out == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null");
        *out = upload_event(*det, meas, seeds, n_events);
    });
}

ckf_status ckf_device_event_destroy(ckf_device_event* ev) {
    return guarded([&] { delete ev; });
}

ckf_status ckf_find_tracks(ckf_algorithm* alg, const ckf_device_detector* det,
                           const ckf_device_event* ev, ckf_device_tracks** out) {
    return guarded([&] {
        if (alg == nullptr || det == nullptr || ev == nullptr || out == nullptr) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        if (ev->n_surfaces != det->h_sfs.size()) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "event was uploaded for another detector");
        }
        *out = find_tracks(*alg, *det, *ev);
    });
}

ckf_status ckf_algorithm_set_scorer(ckf_algorithm* alg, const ckf_scorer* scorer, float threshold) {
    return guarded([&] {
        if (alg == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null algorithm");
        if (!(threshold == threshold)) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "threshold is NaN");
        alg->scorer = scorer;
        alg->score_threshold = threshold;
    });
}

ckf_status ckf_find_tracks_logged(ckf_algorithm* alg, const ckf_device_detector* det,
                                  const ckf_device_event* ev, float* d_features, uint8_t* d_labels,
                                  uint32_t capacity, uint32_t* n_samples, ckf_device_tracks** out) {
    return guarded([&] {
        if (alg == nullptr || det == nullptr || ev == nullptr || out == nullptr ||
            n_samples == nullptr || (capacity > 0u && (d_features == nullptr || d_labels == nullptr))) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        if (ev->n_surfaces != det->h_sfs.size()) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "event was uploaded for another detector");
        }
        if (use_pkf(alg->cfg) || alg->scorer != nullptr) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT,
                            "branch logging runs the plain CKF (no progressive filter, no scorer)");
        }
        struct logging_scope {
            ckf_algorithm& A;
            ~logging_scope() { A.logging = false; }
        } scope{*alg};
        alg->logging = true;
        ckf_device_tracks* T = find_tracks(*alg, *det, *ev);
        try {
            ckf_algorithm& A = *alg;
            const uint32_t n = static_cast<uint32_t>(A.stats.n_links);
            const tamd::ckf_params P = make_params(A, *det, *ev);
            const uint32_t blocks = (n + 1023u) / 1024u;
            std::vector<uint32_t> counts(blocks + 1u, 0u);
            uint32_t total = 0u;
            if (blocks > 0u) {
                uint32_t* d_cnt = nullptr;
                hip_check(hipMalloc(&d_cnt, 4u * (blocks + 1u)), "log buffers");
                try {
                    hip_check(tamd::launch_log_count(P, n, d_cnt, A.stream), "log count");
                    hip_check(hipMemcpyAsync(counts.data(), d_cnt, 4u * blocks, hipMemcpyDeviceToHost,
                                             A.stream), "log count");
                    hip_check(hipStreamSynchronize(A.stream), "log count");
                    for (uint32_t b = 0; b < blocks; ++b) {
                        const uint32_t c = counts[b];
                        counts[b] = total;
                        total += c;
                    }
                    hip_check(hipMemcpyAsync(d_cnt, counts.data(), 4u * blocks, hipMemcpyHostToDevice,
                                             A.stream), "log offsets");
                    if (capacity > 0u) {
                        hip_check(tamd::launch_log_write(P, n, d_cnt, d_features, d_labels, capacity,
                                                         A.stream), "log write");
                    }
                    hip_check(hipStreamSynchronize(A.stream), "log write");
                } catch (...) {
                    (void)hipFree(d_cnt);
                    throw;
                }
                (void)hipFree(d_cnt);
            }
            *n_samples = total;
        } catch (...) {
            delete T;
            throw;
        }
        *out = T;
    });
}

ckf_status ckf_find_tracks_host(ckf_algorithm* alg, const ckf_detector* det,
                                const ckf_bfield* field, const ckf_measurements* meas,
                                const ckf_bound_params* seeds, ckf_device_tracks** out) {
    return guarded([&] {
        if (alg == nullptr || out == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null");
        ckf_device_detector* D = upload_detector(det, field);
        ckf_device_event* E = nullptr;
        try {
            E = upload_event(*D, meas, seeds, 1u);
            *out = find_tracks(*alg, *D, *E);
        } catch (...) {
            delete E;
            delete D;
            throw;
        }
        delete E;
        delete D;
    });
}

ckf_status ckf_tracks_size(const ckf_device_tracks* trk, uint32_t* n_tracks,
                           uint32_t* n_constituents) {
    return guarded([&] {
        if (trk == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null tracks");
        if (n_tracks) *n_tracks = trk->n_tracks;
        if (n_constituents) *n_constituents = trk->n_const;
    });
}

ckf_status ckf_tracks_event_offsets(const ckf_device_tracks* trk, uint32_t* n_events,
                                    uint32_t* track_offsets, uint32_t* constituent_offsets) {
    return guarded([&] {
        if (trk == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null tracks");
        if (n_events) *n_events = trk->n_events;
        if (track_offsets) {
            std::memcpy(track_offsets, trk->ev_track_off.data(), 4u * trk->ev_track_off.size());
        }
        if (constituent_offsets) {
            std::memcpy(constituent_offsets, trk->ev_const_off.data(), 4u * trk->ev_const_off.size());
        }
    });
}

ckf_status ckf_tracks_download(const ckf_device_tracks* trk, ckf_track_candidates* out) {
    return guarded([&] {
        if (trk == nullptr || out == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null");
        const size_t nt = trk->n_tracks, nc = trk->n_const;
        out->n_tracks = trk->n_tracks;
        out->n_constituents = trk->n_const;
        auto d2h = [](void* dst, const void* src, size_t bytes) {
            if (dst != nullptr && bytes > 0) {
                hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "download");
            }
        };
        d2h(out->fit_outcome, trk->outcome, 2u * nt);
        d2h(out->params_vector, trk->vec, 24u * nt);
        d2h(out->params_covariance, trk->cov, 144u * nt);
        d2h(out->ndf, trk->ndf, 4u * nt);
        d2h(out->chi2, trk->chi2, 4u * nt);
        d2h(out->pval, trk->pval, 4u * nt);
        d2h(out->nholes, trk->nholes, 4u * nt);
        d2h(out->constituent_index, trk->cons, 4u * nc);
        if (out->constituent_offsets != nullptr) {
            // offsets index the whole (all-event) constituent array
            d2h(out->constituent_offsets, trk->offsets, 4u * (nt + 1u));
        }
        if (out->params_surface_link != nullptr && nt > 0) {
            std::vector<uint32_t> sf(nt);
            d2h(sf.data(), trk->surface, 4u * nt);
            for (size_t i = 0; i < nt; ++i) {
                // an empty progressive-filter entry has no surface
                out->params_surface_link[i] = sf[i] < trk->ids.size() ? trk->ids[sf[i]] : ~0ull;
            }
        }
        if (out->constituent_type != nullptr) {
            std::fill(out->constituent_type, out->constituent_type + nc, uint16_t(0));
        }
    });
}

ckf_status ckf_tracks_destroy(ckf_device_tracks* trk) {
    return guarded([&] { delete trk; });
}


}  // extern "C"

namespace {
// scoped device copy of a host array; also the destination of downloads
template <typename T>
struct scoped_dev {
    T* p = nullptr;
    size_t n = 0;
    scoped_dev(const T* src, size_t count) : n(count) {
        p = dev_copy(src, count);
    }
    explicit scoped_dev(size_t count) : n(count) { p = dev_copy<T>(nullptr, count); }
    ~scoped_dev() {
        if (p) (void)hipFree(p);
    }
    void download(T* dst) const {
        if (dst != nullptr && n > 0) {
            hip_check(hipMemcpy(dst, p, n * sizeof(T), hipMemcpyDeviceToHost), "download");
        }
    }
};
}  // namespace

extern "C" {

ckf_status ckf_debug_kf_update(uint32_t n, const float* vec, const float* cov, const float* meas,
                               const float* var, const uint32_t* dims, const uint8_t* subspace,
                               float* chi2, int32_t* status, float* out_vec, float* out_cov) {
    return guarded([&] {
        if (n == 0u) return;
        if (!vec || !cov || !meas || !var || !dims || !subspace || !chi2 || !status || !out_vec ||
            !out_cov) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        std::vector<float2> pos(n), vv(n);
        std::vector<uint32_t> info(n);
        for (uint32_t i = 0; i < n; ++i) {
            pos[i] = make_float2(meas[2u * i], meas[2u * i + 1u]);
            vv[i] = make_float2(var[2u * i], var[2u * i + 1u]);
            if ((dims[i] != 1u && dims[i] != 2u) || subspace[2u * i] > 1u ||
                (dims[i] == 2u && (subspace[2u * i + 1u] > 1u ||
                                   subspace[2u * i + 1u] == subspace[2u * i]))) {
                throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "bad measurement");
            }
            info[i] = dims[i] | (uint32_t(subspace[2u * i]) << 8) | (uint32_t(subspace[2u * i + 1u]) << 16);
        }
        scoped_dev<float> dv(vec, 6u * size_t(n)), dc(cov, 36u * size_t(n));
        scoped_dev<float2> dp(pos.data(), n), dva(vv.data(), n);
        scoped_dev<uint32_t> di(info.data(), n);
        scoped_dev<float> dchi(n), dov(6u * size_t(n)), doc(36u * size_t(n));
        scoped_dev<int32_t> dst(n);
        hip_check(tamd::launch_debug_kf(n, dv.p, dc.p, dp.p, dva.p, di.p, dchi.p, dst.p, dov.p,
                                        doc.p, nullptr),
                  "debug kf");
        hip_check(hipDeviceSynchronize(), "sync");
        dchi.download(chi2);
        dst.download(status);
        dov.download(out_vec);
        doc.download(out_cov);
    });
}

ckf_status ckf_debug_propagate(const ckf_finding_config* cfg, const ckf_detector* det,
                               const ckf_bfield* field, uint32_t n, const uint32_t* start,
                               const float* vec, const float* cov, int apply_start_material,
                               int32_t* status, uint32_t* end_surface, float* out_vec,
                               float* out_cov, float* path, uint32_t* n_steps) {
    return guarded([&] {
        if (cfg == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null config");
        validate_config(*cfg);
        if (n == 0u) return;
        if (!start || !vec || !cov || !status || !end_surface || !out_vec || !out_cov || !path ||
            !n_steps) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        ckf_device_detector* D = upload_detector(det, field);
        try {
            for (uint32_t i = 0; i < n; ++i) {
                if (start[i] >= D->h_sfs.size()) {
                    throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "start surface out of range");
                }
            }
            ckf_algorithm A;
            A.cfg = *cfg;
            ckf_device_event E;
            tamd::ckf_params P = make_params(A, *D, E);
            scoped_dev<uint32_t> ds(start, n);
            scoped_dev<float> dv(vec, 6u * size_t(n)), dc(cov, 36u * size_t(n));
            scoped_dev<int32_t> dst(n);
            scoped_dev<uint32_t> des(n), dns(n);
            scoped_dev<float> dov(6u * size_t(n)), doc(36u * size_t(n)), dpath(n);
            hip_check(tamd::launch_debug_propagate(P, n, ds.p, dv.p, dc.p, apply_start_material,
                                                   dst.p, des.p, dov.p, doc.p, dpath.p, dns.p,
                                                   nullptr),
                      "debug propagate");
            hip_check(hipDeviceSynchronize(), "sync");
            dst.download(status);
            des.download(end_surface);
            dov.download(out_vec);
            doc.download(out_cov);
            dpath.download(path);
            dns.download(n_steps);
        } catch (...) {
            delete D;
            throw;
        }
        delete D;
    });
}


ckf_status ckf_debug_set_kernel_timing(ckf_algorithm* alg, int enable) {
    return guarded([&] {
        if (alg == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null algorithm");
        alg->timing = enable != 0;
    });
}

ckf_status ckf_debug_set_schedule(ckf_algorithm* alg, uint32_t lanes_per_link) {
    return guarded([&] {
        if (alg == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null algorithm");
        if (lanes_per_link != 0u && lanes_per_link != 1u && lanes_per_link != 2u &&
            lanes_per_link != 4u && lanes_per_link != 8u && lanes_per_link != 16u) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "lanes per link must be 0, 1, 2, 4, 8 or 16");
        }
        alg->schedule_override = lanes_per_link;
    });
}

ckf_status ckf_debug_get_schedule(const ckf_algorithm* alg, uint32_t* lanes_per_link) {
    return guarded([&] {
        if (alg == nullptr || lanes_per_link == nullptr) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        *lanes_per_link = alg->prop_team;
    });
}

ckf_status ckf_debug_kernel_times(const ckf_algorithm* alg, double* ms_find, double* ms_dedup,
                                  double* ms_propagate, double* ms_other, uint32_t* n_steps) {
    return guarded([&] {
        if (alg == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null algorithm");
        if (ms_find) *ms_find = alg->t_find;
        if (ms_dedup) *ms_dedup = alg->t_dedup;
        if (ms_propagate) *ms_propagate = alg->t_prop;
        if (ms_other) *ms_other = alg->t_other;
        if (n_steps) *n_steps = alg->n_prop_launches;
    });
}

ckf_status ckf_release_cached_memory(void) {
    return guarded([&] { result_cache().release(); });
}

ckf_status ckf_debug_set_verdict_log(ckf_algorithm* alg, uint32_t capacity) {
    return guarded([&] {
        if (alg == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null algorithm");
        alg->verdict_cap = capacity;
        if (capacity == 0u) {
            alg->verdict_log.release();
            alg->verdict_n.release();
        }
    });
}

ckf_status ckf_debug_get_verdicts(const ckf_algorithm* alg, uint64_t* keys, uint32_t capacity,
                                  uint32_t* n_records) {
    return guarded([&] {
        if (alg == nullptr || n_records == nullptr) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        }
        *n_records = 0u;
        if (alg->verdict_cap == 0u || alg->verdict_n.p == nullptr) return;
        uint32_t n = 0u;
        hip_check(hipStreamSynchronize(alg->stream), "sync");
        hip_check(hipMemcpy(&n, alg->verdict_n.p, 4u, hipMemcpyDeviceToHost), "verdict count");
        *n_records = n;
        const uint32_t m = std::min({n, capacity, alg->verdict_cap});
        if (m > 0u && keys != nullptr) {
            hip_check(hipMemcpy(keys, alg->verdict_log.p, 8u * size_t(m), hipMemcpyDeviceToHost),
                      "verdict download");
        }
    });
}

ckf_status ckf_debug_propagate_split(const ckf_algorithm* alg, double* ms_propagate_kernel,
                                     double* ms_propagate_emit) {
    return guarded([&] {
        if (alg == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null algorithm");
        if (ms_propagate_kernel) *ms_propagate_kernel = alg->t_prop_work;
        if (ms_propagate_emit) *ms_propagate_emit = alg->t_prop_emit;
    });
}

ckf_status ckf_debug_find_split(const ckf_algorithm* alg, double* ms_find_kernel,
                                double* ms_find_emit) {
    return guarded([&] {
        if (alg == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null algorithm");
        if (ms_find_kernel) *ms_find_kernel = alg->t_find_work;
        if (ms_find_emit) *ms_find_emit = alg->t_find_emit;
    });
}

}  // extern "C"
