/*
 * traccc_amd -- greedy ambiguity resolution (C ABI).
 *
 * Replaces traccc's greedy ambiguity resolution operator, the step after the
 * CKF (and the fit):
 *
 *   traccc::host::greedy_ambiguity_resolution_algorithm::operator()(tracks)
 *     core/include/traccc/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.hpp:44-55
 *     core/src/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.cpp:19-254
 *   traccc::cuda::greedy_ambiguity_resolution_algorithm::operator()(tracks)
 *     device/cuda/src/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.cu:88-711
 *
 * Tracks sharing measurements are removed one at a time, worst first: the
 * worst track has the highest fraction of shared measurements, then the
 * lowest p-value; a removal can make another track's measurements unshared.
 * The iteration ends when no accepted track has max_shared_meas shared
 * measurements, or after max_iterations removals.  Tracks with fewer than
 * min_meas_per_track measurements are dropped before it starts.
 *
 * The result is the reference host algorithm's: the same set of tracks kept,
 * reported in input order (the host's order; the CUDA reference returns them
 * in its internal sort order).  Exact ties in (shared fraction, p-value) go
 * against the higher track index (the reference leaves their order to
 * std::sort); a NaN p-value orders as the lowest.
 *
 * Same conventions as ckf.h: plain C, ckf_status returns, thread-local
 * message in ckf_last_error().
 */
#ifndef TRACCC_AMD_AMBIGUITY_H
#define TRACCC_AMD_AMBIGUITY_H

#include <stdint.h>

#include "traccc_amd/ckf.h"
#include "traccc_amd/fitting.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of traccc::ambiguity_resolution_config
 * (core/include/traccc/ambiguity_resolution/ambiguity_resolution_config.hpp:16-27). */
typedef struct ckf_ambiguity_config {
    uint32_t min_meas_per_track; /* 3 */
    uint32_t max_iterations;     /* UINT32_MAX */
    uint32_t max_shared_meas;    /* 1; must be > 0 (the reference asserts it) */
} ckf_ambiguity_config;

void ckf_ambiguity_config_default(ckf_ambiguity_config* cfg);

/* The tracks to resolve, as the reference reads them from its
 * track_container: each track's p-value and the identifiers of its
 * measurement constituents (greedy_ambiguity_resolution_algorithm.cpp:48-73).
 * Track t owns measurement_id[constituent_offsets[t] .. constituent_offsets[t+1]).
 * Every identifier must be < id_bound (the CUDA reference requires the
 * identifiers to be exactly 0 .. n_measurements - 1). */
typedef struct ckf_resolver_tracks {
    uint32_t n_tracks;
    const uint32_t* constituent_offsets; /* [n_tracks + 1], non-decreasing, from 0 */
    const uint32_t* measurement_id;      /* [constituent_offsets[n_tracks]] */
    const float* pval;                   /* [n_tracks] */
    uint32_t id_bound;
} ckf_resolver_tracks;

/* Statistics of the last call on a resolver. */
typedef struct ckf_resolver_stats {
    uint32_t n_tracks;      /* input tracks                                      */
    uint32_t n_candidates;  /* tracks with at least min_meas_per_track measurements */
    uint32_t n_removed;     /* tracks the iterations removed                     */
    uint32_t n_accepted;    /* tracks kept                                       */
    uint32_t n_rounds;      /* parallel removal rounds (0 on the sequential path) */
    uint32_t sequential;    /* 1 when the one-at-a-time path ran                 */
} ckf_resolver_stats;

typedef struct ckf_resolver ckf_resolver;

/* traccc: greedy_ambiguity_resolution_algorithm(config, mr[, copy, stream]).
 * max_shared_meas == 0 gives CKF_ERR_INVALID_ARGUMENT. */
ckf_status ckf_resolver_create(const ckf_ambiguity_config* cfg, ckf_resolver** out);
ckf_status ckf_resolver_destroy(ckf_resolver* r);
/* hipStream_t as void*; NULL = the default stream. */
ckf_status ckf_resolver_set_stream(ckf_resolver* r, void* hip_stream);
ckf_status ckf_resolver_last_stats(const ckf_resolver* r, ckf_resolver_stats* stats);

/* The operator on host arrays: accepted[t] = 1 for every track kept, 0
 * otherwise (caller-owned, n_tracks bytes); *n_accepted = their number.
 * Inputs are copied to HBM; returns when the result is on the host. */
ckf_status ckf_resolve_ambiguities(ckf_resolver* r, const ckf_resolver_tracks* tracks,
                                   uint8_t* accepted, uint32_t* n_accepted);

/* The same on arrays already in HBM (tracks' pointers are device
 * pointers); d_accepted is device memory of n_tracks bytes.  The launches
 * go to the resolver's stream; the call returns when d_accepted is written. */
ckf_status ckf_resolve_ambiguities_device(ckf_resolver* r, const ckf_resolver_tracks* tracks,
                                          uint8_t* d_accepted, uint32_t* n_accepted);

/* The operator on fitted tracks (ckf_fit_tracks of the events `ev`): each
 * fitted track's p-value and its candidate's measurements, identified by
 * their index in the batch (event offset + index in the event; the
 * reference's identifiers are these when they number the measurements).
 * accepted: caller-owned host array of ckf_fitted_size() tracks. */
ckf_status ckf_resolve_fitted(ckf_resolver* r, const ckf_device_event* ev,
                              const ckf_device_fitted* fitted, uint8_t* accepted,
                              uint32_t* n_accepted);

#ifdef __cplusplus
}
#endif

#endif /* TRACCC_AMD_AMBIGUITY_H */
