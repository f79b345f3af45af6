// Device-side data layout of the CKF step pipeline (shared by the kernels and
// the host runtime).
//
// HBM layout (structure of arrays, one lane per branch):
//   bound state : 28 x 4 B columns of `cap` rows: 6 parameters, the 21-entry
//                 upper triangle of the covariance, the surface index (u32)
//   link        : 32 B record {prev, meas, seed, step|skipped|consecutive,
//                 chi2, chi2_sum, ndf_sum, state_ref}
//   measurement : float2 position, float2 variance, u32 {dims, subspace}
#pragma once

#include <cstdint>

#include "../core/propagator.hpp"
#include "../scorer/mlp_view.hpp"

namespace tamd {

constexpr int kStateCols = 28;  // 6 + 21 + 1
constexpr int kTickets = 9;     // work queues of the control block
constexpr int kLookbacks = 4;   // compactions with their own look-back words:
                                // find emit, dedup, propagate emit, build
constexpr uint32_t kInvalid = 0xffffffffu;

// state_ref encoding: where a link's (filtered) state lives
constexpr uint32_t kRefKindShift = 30;
constexpr uint32_t kRefFiltered = 0u;  // filtered-state slot
constexpr uint32_t kRefInput = 1u;     // previous step's propagated state slot
constexpr uint32_t kRefSeed = 2u;      // seed index (step 0)
constexpr uint32_t kRefMask = (1u << kRefKindShift) - 1u;

struct dlink {
    uint32_t prev;      // global index of the link it continues, kInvalid at step 0
    uint32_t meas;      // global measurement index, kInvalid for a hole
    uint32_t seed;      // global seed index
    uint32_t packed;    // step (16 b) | n_skipped (8 b) << 16 | n_consecutive (8 b) << 24
    float chi2;
    float chi2_sum;
    uint32_t ndf_sum;
    uint32_t state_ref;
};
static_assert(sizeof(dlink) == 32, "link record is 32 B");

// What k_find hands k_find_emit so the links are built without re-reading
// the inputs: per input its previous link, seed, hole counts, running sums
// and own state reference (written before the updates, so none of it stays
// in registers through them); per filtered slot the candidate.
struct stage_head {
    uint32_t prev;      // global index of the input's link, kInvalid at step 0
    uint32_t seed;
    uint32_t hist;      // n_skipped | n_consecutive << 8
    uint32_t hole_ref;  // state_ref of a hole link (the input's own state)
    float chi2_sum;     // the input's running sums
    uint32_t ndf_sum;
    uint32_t pad[2];
};
struct stage_cand {
    uint32_t meas;      // measurement | (dims == 2) << 31; kInvalid: the input's hole
    float chi2;
};
static_assert(sizeof(stage_head) == 32 && sizeof(stage_cand) == 8, "stage record sizes");

TAMD_HD uint32_t link_step(const dlink& l) { return l.packed & 0xffffu; }
TAMD_HD uint32_t link_skipped(const dlink& l) { return (l.packed >> 16) & 0xffu; }
TAMD_HD uint32_t link_consec(const dlink& l) { return (l.packed >> 24) & 0xffu; }
TAMD_HD uint32_t pack_link(uint32_t step, uint32_t skipped, uint32_t consec) {
    return (step & 0xffffu) | ((skipped > 255u ? 255u : skipped) << 16) |
           ((consec > 255u ? 255u : consec) << 24);
}

// A counter alone on its 128-byte cache line.  The work-queue counters take
// an atomic per partition from every wave of a grid; counters sharing a line
// (with each other, or with the fields every wave reads at kernel start) are
// served one after another by that line's L2 channel.
struct alignas(128) line_u32 {
    uint32_t v;
};
struct alignas(128) line_u64 {
    uint64_t v;
};
static_assert(sizeof(line_u32) == 128 && sizeof(line_u64) == 128, "one counter per line");

// Step control block in device memory; the step loop reads and advances it
// without host round trips.  The fields every wave reads at kernel start
// share the first line; each counter that waves hammer with atomics has its
// own.
// A deferred propagation: link, visited layers (2 words), RK steps, path,
// the free state (11 words), the transport Jacobian (31 words), state ref.
constexpr uint32_t kDeferWords = 48u;

struct step_ctl {
    uint32_t step;        // current CKF step
    uint32_t epoch;       // tag of this step's look-back descriptors
    uint32_t n_in;        // inputs of this step
    uint32_t link_base;   // global index of this step's first link
    uint32_t n_links;     // links made this step (written by find)
    uint32_t n_live;      // inputs of the next step (written by propagate)
    uint32_t n_tips;      // tips recorded before this step
    uint32_t n_tips_step; // tips recorded this step (written by propagate)
    uint32_t overflow;    // bit 0 links, bit 1 per-step states, bit 2 filtered, bit 3 tips, bit 4 inputs
    uint32_t done;
    uint32_t n_steps_run;
    uint32_t n_tracks;
    uint32_t n_constituents;
    uint32_t prev_link_base;  // global index of the previous step's first link
    uint64_t n_links_total;
    uint64_t n_scored;        // scorer rows of the finished steps
    line_u32 ticket[kTickets];  // work-queue tickets: [0] progressive filter, [1] votes /
                                // progressive filter, [2] free, [3] build, [4] find emit,
                                // [5] propagate emit, [6] dedup emit, [7] free, [8] free
    line_u32 exit_count;  // blocks of k_prop_emit that have finished this step
    line_u32 n_rows;      // scorer rows scored this step (added up by k_score_update)
    line_u64 n_props;     // propagations so far
    line_u32 n_defer;     // links whose propagation k_propagate_resume continues (this step)
    line_u32 xcd_ticket[4][8];  // per-XCD work queues (one counter per XCD): find, scorer
                                // list, propagate, dedup mark
};

struct ckf_params {
    // detector + field
    const ckf_surface* sfs;
    const ckf_layer* layers;
    uint32_t n_layers;
    const layer_group* groups;  // build_layer_groups of the layers
    uint32_t n_groups;
    // propagations deferred at their second layer search: kDeferWords
    // words per entry, word w of entry k at defer[w * defer_cap + k]
    uint32_t* defer;
    uint32_t defer_cap;
    uint32_t n_surfaces;
    field_view<float> field;
    nav_config<float> nav;
    // finding config
    uint32_t n_branch_surface;  // max_num_branches_per_surface
    uint32_t n_branch_seed;     // max_num_branches_per_seed
    uint32_t min_cands, max_cands;
    uint32_t max_skip, max_consec;
    uint32_t dup_min_len;
    float chi2_max;
    // inputs (batch of events)
    uint32_t n_events;
    uint32_t n_seeds;
    uint32_t n_meas;
    const float2* meas_pos;
    const float2* meas_var;
    const uint32_t* meas_info;    // dims | subspace0 << 8 | subspace1 << 16
    const uint32_t* meas_ranges;  // [n_events][n_surfaces + 1]: event offset, then upper bounds
    const uint32_t* seed_event;
    const uint32_t* seed_surface;
    const float* seed_vec;        // [n_seeds][6]
    const float* seed_cov;        // [n_seeds][36]
    // work buffers
    float* state[2];              // ping-pong propagated states, SoA [28][state_cap]
    uint32_t state_cap;           // links per step
    float* flt;                   // filtered states SoA [28][flt_cap], slot = k * in_cap + input
    stage_head* stage_in;         // [in_cap] the input's link bookkeeping, from k_find
    stage_cand* stage_slot;       // [flt_cap] the filtered slot's candidate (slot k = 0 of
                                  // an input holds the hole when nothing was kept)
    uint32_t flt_cap;
    uint32_t in_cap;              // inputs per step
    dlink* links;
    uint32_t link_cap;
    uint32_t* live;               // [in_cap] previous-step link slots of this step's inputs
    uint32_t* tips;               // [tip_cap] global link indices
    uint32_t tip_cap;
    uint32_t* seed_first;         // [n_seeds] local index of the seed's first link this step
    uint32_t* seed_base;          // [n_seeds] alive-prefix at that link (dedup steps)
    uint32_t* alive_prefix;       // [state_cap] bit 31 = alive, low bits = exclusive alive prefix
    uint64_t* dedup_head;         // [n_meas + 1] epoch << 32 | link (epoch-tagged list heads)
    uint32_t* dedup_next;         // [state_cap]
    uint64_t* lb_lo[kLookbacks];  // look-back words per compaction: low counter channel
    uint64_t* lb_hi[kLookbacks];  // and high counter channel
    uint32_t lb_cap;
    step_ctl* ctl;
    // build outputs
    uint32_t* tip_keep;           // [tip_cap] bit 31 kept | n_cands
    uint32_t* tip_rank;           // [tip_cap] kept prefix
    uint32_t* tip_cprefix;        // [tip_cap] constituent prefix
    uint32_t* seg_count;          // [n_events][max_cands] kept tracks per (event, step)
    uint32_t* seg_ccount;         // [n_events][max_cands] constituents per (event, step)
    uint32_t* seg_base;           // [n_events][max_cands] kept prefix at segment head
    uint32_t* seg_cbase;          // [n_events][max_cands] constituent prefix at segment head
    uint32_t* seg_off;            // [n_events][max_cands] output track offset
    uint32_t* seg_coff;           // [n_events][max_cands] output constituent offset
    uint32_t* event_track_off;    // [n_events + 1]
    uint32_t* event_const_off;    // [n_events + 1]
    const uint32_t* event_meas_off;  // [n_events]
    const float* event_win_var;      // [n_events] largest loc0 variance, -1 = no window search
    // output arrays
    uint16_t* out_outcome;
    uint32_t* out_surface;        // surface index of the seed
    float* out_vec;               // [n][6]
    float* out_cov;               // [n][36]
    float* out_ndf;
    float* out_chi2;
    float* out_pval;
    uint32_t* out_nholes;
    uint32_t* out_offsets;        // [n + 1]
    uint32_t* out_const;          // [n_const]
    uint32_t out_cap;
    uint32_t out_const_cap;
    uint64_t* vote_key;           // [n]; written when measurement votes run, else null
    // branch logging (training data of the scorer): features of every
    // filtered slot, copied to every measurement link; survive marks by build
    float* feat_slot;             // [flt_cap][12] or null
    float* log_feat;              // [link_cap][12] or null
    uint8_t* log_label;           // [link_cap] or null
    uint64_t* seed_best;          // [n_seeds] best kept tip of each seed (logging)
    // scorer-pruned finding: per input the list_len best candidates in
    // (chi2, index) order with their features, scored before the updates
    uint32_t scored;              // 1: k_find takes its candidates from the list
    uint32_t list_len;
    uint32_t* list_m;             // [list_len][in_cap] measurement
    float* list_c;                // [list_len][in_cap] chi2
    uint32_t* find_count;         // [in_cap] branches each input keeps this step
    uint8_t* prop_flag;           // [state_cap] propagation outcome: bit 0 next input, bits 1-2 tips
    uint64_t* seed_run_min;       // [seeds] scorer feature 9: lowest rank-0 running chi2/ndf
    uint32_t* list_seed;          // [in_cap] seed of each listed input
    uint32_t* list_n;             // [in_cap] entries
    short* list_row;              // [list_len][in_cap][16] each listed candidate's MLP row:
                                  // features normalised to BF16, feature 9 raw FP32 in 12-13
    mlp_dev scorer;
    float score_threshold;        // logit below which a candidate is pruned
    // candidate-verdict log (diagnostics): one key per scored candidate,
    // step << 48 | input << 24 | measurement, bit 63 = kept
    uint64_t* verdict_log;        // [verdict_cap] or null
    uint32_t* verdict_n;          // records written (may exceed verdict_cap)
    uint32_t verdict_cap;
};

}  // namespace tamd
