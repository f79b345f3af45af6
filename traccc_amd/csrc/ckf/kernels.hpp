// Host-side launchers of the CKF kernels (defined in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_types.hpp"

namespace tamd {

struct launch_shape {
    uint32_t find_blocks;
    uint32_t dedup_blocks;
    uint32_t prop_blocks;
    uint32_t build_blocks;
    uint32_t prop_team;  // lanes per link in k_propagate: 1, 2, 4, 8 or 16
    uint32_t list_blocks;   // scorer-pruned finding: candidate lists
    uint32_t score_blocks;  // and the MLP + updates over them (64 inputs per wave)
    uint32_t emit_blocks;   // the compaction kernels (k_*_emit)
    uint32_t resume_blocks; // deferred propagations (k_propagate_resume)
};

// One CKF step: find_tracks, duplicate removal (active from
// duplicate_removal_minimum_length on), propagation to the next surface with
// the step advance folded into the last block of its compaction kernel.  Every count is read from the
// device step control block, so the sequence can be captured in a graph.
// with_dedup = false drops the duplicate-removal kernels (steps before
// duplicate_removal_minimum_length, where they would exit at once).
hipError_t launch_ckf_step(const ckf_params& P, const launch_shape& s, hipStream_t stream,
                           bool with_dedup);
// the same step, kernel group by kernel group (timed runs)
// launch_find = the find kernel (plain or logged; scorer-pruned: the
// candidate lists, then the MLP and the updates), then its link compaction
hipError_t launch_find(const ckf_params& P, const launch_shape& s, hipStream_t stream);
hipError_t launch_find_work(const ckf_params& P, const launch_shape& s, hipStream_t stream);
hipError_t launch_find_emit(const ckf_params& P, const launch_shape& s, hipStream_t stream);
hipError_t launch_dedup(const ckf_params& P, const launch_shape& s, hipStream_t stream);
hipError_t launch_propagate(const ckf_params& P, const launch_shape& s, hipStream_t stream,
                            bool fused_advance);
// launch_propagate = the propagation kernel (at the schedule's team size),
// then its compaction (k_prop_emit)
hipError_t launch_propagate_work(const ckf_params& P, const launch_shape& s, hipStream_t stream);
hipError_t launch_prop_emit(const ckf_params& P, const launch_shape& s, hipStream_t stream,
                            bool fused_advance);
hipError_t launch_advance(const ckf_params& P, hipStream_t stream);

// build_tracks: tip scan, (event, step) segment table, output fill.
hipError_t launch_build_count(const ckf_params& P, const launch_shape& s, hipStream_t stream);
// one lane per tip (n_tips from the control block after launch_build_count)
hipError_t launch_build_write(const ckf_params& P, const launch_shape& s, uint32_t n_tips,
                              hipStream_t stream);

// Branch log (scorer training data): measurement links [0, n) in order into
// caller arrays; block_n / block_off hold one count / offset per 1024 links.
hipError_t launch_label_best(const ckf_params& P, const launch_shape& s, hipStream_t stream);
hipError_t launch_log_count(const ckf_params& P, uint32_t n, uint32_t* block_n, hipStream_t stream);
hipError_t launch_log_write(const ckf_params& P, uint32_t n, const uint32_t* block_off,
                            float* features, uint8_t* labels, uint32_t cap, hipStream_t stream);

// Measurement-vote pruning of built tracks (max_num_tracks_per_measurement > 0).
struct vote_params {
    step_ctl* ctl;                  // epoch and ticket[1] of the compaction scan
    uint32_t n_tracks, n_events, K;
    float min_fraction;
    const uint32_t* track_off;      // [n_events + 1] built tracks per event
    const uint32_t* event_meas_off; // [n_events]
    const uint32_t* offsets;        // [n_tracks + 1]
    const uint32_t* cons;           // event-relative measurement indices
    const uint64_t* key;            // [n_tracks] (p-value bits << 32) | ~track
    uint64_t* slots;                // [n_meas_total][K], zeroed
    uint32_t* keep;                 // [n_tracks]
    uint32_t* rank;                 // [n_tracks]
    uint32_t* cpos;                 // [n_tracks]
    uint32_t* totals;               // [2] kept tracks, kept constituents
    uint32_t* new_track_off;        // [n_events + 1]
    uint32_t* new_const_off;        // [n_events + 1]
    uint64_t* lb_lo;  // look-back words (low / high counter channel)
    uint64_t* lb_hi;
    // built track columns
    const uint16_t* outcome;
    const uint32_t* surface;
    const float* vec;
    const float* cov;
    const float* ndf;
    const float* chi2;
    const float* pval;
    const uint32_t* nholes;
    // kept track columns
    uint16_t* o_outcome;
    uint32_t* o_surface;
    float* o_vec;
    float* o_cov;
    float* o_ndf;
    float* o_chi2;
    float* o_pval;
    uint32_t* o_nholes;
    uint32_t* o_offsets;
    uint32_t* o_cons;
};
hipError_t launch_votes(const vote_params& V, uint32_t n_cu, hipStream_t stream);

// Progressive Kalman filter: per-seed walk + constituent scan, then the
// output fill (after the host has sized the result from ctl->n_constituents).
struct pkf_params {
    uint32_t* states;    // [n_seeds][max_states] global measurement indices
    uint32_t* n_states;  // [n_seeds]; 0 = empty entry
    float* chi2;         // [n_seeds]
    uint32_t* ndf;       // [n_seeds] sum of measurement dimensions
    uint32_t* holes;     // [n_seeds]
    uint32_t* cpos;      // [n_seeds] constituent offset
    uint32_t max_states;
    uint64_t* lb_lo;  // look-back words (low / high counter channel)
    uint64_t* lb_hi;
};
hipError_t launch_pkf(const ckf_params& P, const pkf_params& Q, uint32_t n_cu, hipStream_t stream);
hipError_t launch_pkf_write(const ckf_params& P, const pkf_params& Q, uint32_t n_cu,
                            hipStream_t stream);

// diagnostics (include/traccc_amd/ckf_debug.h)
hipError_t launch_debug_kf(uint32_t n, const float* vec, const float* cov, const float2* meas,
                           const float2* var, const uint32_t* info, float* chi2, int32_t* status,
                           float* out_vec, float* out_cov, hipStream_t stream);
hipError_t launch_debug_propagate(const ckf_params& P, uint32_t n, const uint32_t* start,
                                  const float* vec, const float* cov, int apply_start_material,
                                  int32_t* status, uint32_t* end_surface, float* out_vec,
                                  float* out_cov, float* path, uint32_t* n_steps,
                                  hipStream_t stream);

}  // namespace tamd
