/*
 * traccc_amd -- Kalman track fit with a two-filter smoother (C ABI).
 *
 * Replaces traccc's track-fitting operator on the track candidates the CKF
 * found:
 *
 *   traccc::device::kalman_fitting_algorithm::operator()(det, field, track_candidates)
 *     device/common/include/traccc/fitting/device/kalman_fitting_algorithm.hpp:91-101
 *   traccc::kalman_fitter::fit / filter / smooth
 *     core/include/traccc/fitting/kalman_filter/kalman_fitter.hpp:177-330
 *   traccc::two_filters_smoother
 *     core/include/traccc/fitting/kalman_filter/two_filters_smoother.hpp:38-294
 *
 * Per candidate: the seed parameters (covariance inflated) are filtered
 * forward through every sensitive surface the navigation reaches (material
 * first, then the gain-matrix update on the candidate's measurement of that
 * surface; a surface without one is a hole), then smoothed backward from the
 * last state, surface by surface, combining the backward prediction with the
 * forward filtered state.  Statistics and the fit outcome follow
 * kalman_fitter::update_statistics / check_fitting_result
 * (kalman_fitter.hpp:419-550).
 *
 * Same conventions as ckf.h: plain C, ckf_status returns, thread-local
 * message in ckf_last_error().
 */
#ifndef TRACCC_AMD_FITTING_H
#define TRACCC_AMD_FITTING_H

#include <stdint.h>

#include "traccc_amd/ckf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* traccc track_fit_outcome (core/include/traccc/edm/track_fit_outcome.hpp:16-27) */
typedef enum ckf_fit_outcome {
    CKF_OUTCOME_UNKNOWN = 0,
    CKF_OUTCOME_SUCCESS = 1,
    CKF_OUTCOME_FAILURE_NON_POSITIVE_NDF = 2,
    CKF_OUTCOME_FAILURE_NOT_ALL_FITTED = 3,
    CKF_OUTCOME_FAILURE_NOT_ALL_SMOOTHED = 4,
    CKF_OUTCOME_FAILURE_FITTER = 5,
    CKF_OUTCOME_FAILURE_SMOOTHER = 6,
    CKF_OUTCOME_FAILURE_FORWARD_PROPAGATION = 7,
    CKF_OUTCOME_FAILURE_BACKWARD_PROPAGATION = 8
} ckf_fit_outcome;

/* Track-state flags (traccc::edm::track_state IS_HOLE_MASK / IS_SMOOTHED_MASK,
 * core/include/traccc/edm/track_state_collection.hpp:36-37). */
#define CKF_STATE_HOLE 0x01u
#define CKF_STATE_SMOOTHED 0x02u

/* Mirror of traccc::fitting_config (core/include/traccc/fitting/fitting_config.hpp:20-45). */
typedef struct ckf_fitting_config {
    uint32_t n_iterations;                  /* 1 */
    ckf_propagation_config propagation;     /* stepping + navigation */
    uint32_t max_step_counts;               /* RK steps allowed between two surfaces (100) */
    float min_p;                            /* 0.1 GeV */
    float min_pT;                           /* 0.6 GeV */
    int32_t ptc_pdg;                        /* 211 (pi+) */
    float ptc_mass;                         /* GeV */
    float ptc_charge;                       /* e   */
    float covariance_inflation_factor;      /* 1e3 */
    uint32_t surface_sequence_size_factor;  /* 5: surfaces kept per state for the smoother */
    uint32_t min_surface_sequence_capacity; /* 100 */
    uint32_t do_precise_hole_count;         /* 0 */
} ckf_fitting_config;

typedef struct ckf_fitter ckf_fitter;             /* one configured fitter          */
typedef struct ckf_device_fitted ckf_device_fitted; /* fitted tracks + states in HBM */

/* Fitted tracks (traccc track_collection, track_collection.hpp:183-198) and
 * their track states (track_state_collection.hpp:225-243).  Caller-owned host
 * buffers sized from ckf_fitted_size(); any pointer may be NULL to skip that
 * column.  Track t owns states [state_offsets[t], state_offsets[t+1]) in the
 * candidate's constituent order. */
typedef struct ckf_fitted_tracks {
    uint32_t n_tracks;
    uint32_t n_states;
    uint16_t* fit_outcome;           /* [n_tracks] ckf_fit_outcome              */
    uint64_t* params_surface_link;   /* [n_tracks] smoothed params of the first smoothed state */
    float* params_vector;            /* [n_tracks][6]                           */
    float* params_covariance;        /* [n_tracks][36]                          */
    float* ndf;                      /* [n_tracks] sum of dims of smoothed states - 5 */
    float* chi2;                     /* [n_tracks] sum of backward chi2         */
    float* pval;                     /* [n_tracks]                              */
    uint32_t* nholes;                /* [n_tracks]                              */
    uint32_t* state_offsets;         /* [n_tracks + 1]                          */
    uint8_t* state_flags;            /* [n_states] CKF_STATE_*                  */
    float* filtered_chi2;            /* [n_states]                              */
    float* smoothed_chi2;            /* [n_states]                              */
    float* backward_chi2;            /* [n_states]                              */
    uint64_t* state_surface_link;    /* [n_states]                              */
    float* filtered_vector;          /* [n_states][6]                           */
    float* filtered_covariance;      /* [n_states][36]                          */
    float* smoothed_vector;          /* [n_states][6]                           */
    float* smoothed_covariance;      /* [n_states][36]                          */
    uint32_t* measurement_index;     /* [n_states] into the caller's measurements */
} ckf_fitted_tracks;

void ckf_fitting_config_default(ckf_fitting_config* cfg);

/* traccc: kalman_fitting_algorithm(config, mr, copy, logger).  Invalid
 * settings return CKF_ERR_INVALID_ARGUMENT. */
ckf_status ckf_fitter_create(const ckf_fitting_config* cfg, ckf_fitter** out);
ckf_status ckf_fitter_destroy(ckf_fitter* f);
/* hipStream_t as void*; NULL = the default stream. */
ckf_status ckf_fitter_set_stream(ckf_fitter* f, void* hip_stream);

/* The operator: fit every candidate of `candidates` (a ckf_find_tracks
 * result on the events `ev`, which must stay alive).  Output in HBM; the call
 * returns when it is complete. */
ckf_status ckf_fit_tracks(ckf_fitter* f, const ckf_device_detector* det, const ckf_device_event* ev,
                          const ckf_device_tracks* candidates, ckf_device_fitted** out);

ckf_status ckf_fitted_size(const ckf_device_fitted* r, uint32_t* n_tracks, uint32_t* n_states);
ckf_status ckf_fitted_download(const ckf_device_fitted* r, ckf_fitted_tracks* out);
ckf_status ckf_fitted_destroy(ckf_device_fitted* r);

#ifdef __cplusplus
}
#endif

#endif /* TRACCC_AMD_FITTING_H */
