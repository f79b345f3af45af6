/*
 * CPU oracle for the CKF track-finding path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker / the timed CPU baseline.  The
 * product (traccc_amd) never links, loads or calls it.
 *
 * It is this project's restatement of traccc's host CKF
 * (core/include/traccc/finding/details/combinatorial_kalman_filter.hpp:59-718),
 * templated on the scalar: the FP64 instantiation is the correctness oracle,
 * the FP32 instantiation is the "reference CPU CKF" timed beside the GPU.
 * It shares with the product only the plain C data layouts of ckf.h.
 */
#ifndef TRACCC_AMD_ORACLE_H
#define TRACCC_AMD_ORACLE_H

#include <stdint.h>

#include "traccc_amd/ckf.h"
#include "traccc_amd/fitting.h"
#include "traccc_amd/seeding.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_tracks {
    uint32_t n_tracks;
    uint32_t n_constituents;
    uint64_t* params_surface_link;
    float* params_vector;      /* [n][6]  */
    float* params_covariance;  /* [n][36] */
    double* ndf;
    double* chi2;
    double* pval;
    uint32_t* nholes;
    uint32_t* constituent_offsets; /* [n + 1] */
    uint32_t* constituent_index;
    /* statistics */
    uint32_t n_steps;
    uint64_t n_links;
    uint64_t n_propagations;
    uint64_t n_tips;
} oracle_tracks;

/* precision: 64 (oracle) or 32 (timed CPU baseline). */
int oracle_ckf_run(const ckf_finding_config* cfg, const ckf_detector* det, const ckf_bfield* field,
                   const ckf_measurements* meas, const ckf_bound_params* seeds, int precision,
                   oracle_tracks* out);
void oracle_tracks_free(oracle_tracks* t);

/* The CKF pruned by given candidate verdicts (the learned branch scorer's,
 * read back from the device): per input only the list_len best candidates by
 * (chi2, index) passing the cut are listed, and only the listed ones whose key
 * is in `kept` (sorted ascending) are updated, in order, until
 * max_num_branches_per_surface succeed.  A key is splitmix64 chained as
 * h = mix(h ^ v) from h = 0 over the seed index, the step, the measurement
 * index of every earlier link of the branch (latest first, holes as
 * 0xffffffff) and the candidate's measurement index, low 63 bits.
 * kept = NULL with n_kept = UINT64_MAX keeps every listed candidate. */
int oracle_ckf_run_masked(const ckf_finding_config* cfg, const ckf_detector* det,
                          const ckf_bfield* field, const ckf_measurements* meas,
                          const ckf_bound_params* seeds, int precision, uint32_t list_len,
                          const uint64_t* kept, uint64_t n_kept, oracle_tracks* out);

/* Events spread over n_threads host threads, one CKF instance per thread
 * (traccc's throughput_mt pattern).  Returns wall seconds in *seconds. */
int oracle_ckf_run_many(const ckf_finding_config* cfg, const ckf_detector* det,
                        const ckf_bfield* field, const ckf_measurements* meas,
                        const ckf_bound_params* seeds, uint32_t n_events, int precision,
                        uint32_t n_threads, uint64_t* n_tracks_total, double* seconds);

/* ---- Kalman fit + two-filter smoother (kalman_fitting_algorithm) ---- */
typedef struct oracle_fitted {
    uint32_t n_tracks;
    uint32_t n_states;
    uint16_t* outcome;           /* ckf_fit_outcome */
    uint32_t* params_surface;    /* surface index */
    double* params_vector;       /* [n][6]  */
    double* params_covariance;   /* [n][36] */
    double* ndf;
    double* chi2;
    double* pval;
    uint32_t* nholes;
    uint32_t* state_offsets;     /* [n + 1] */
    uint8_t* state_flags;
    double* filtered_chi2;
    double* smoothed_chi2;
    double* backward_chi2;
    double* filtered_vector;     /* [n_states][6]  */
    double* filtered_covariance; /* [n_states][36] */
    double* smoothed_vector;
    double* smoothed_covariance;
} oracle_fitted;

/* Fit n_tracks candidates of one event: seed parameters per track and the
 * candidate's measurement indices [offsets[t], offsets[t+1]) into `meas`.
 * precision 64 (oracle) or 32 (timed CPU baseline); tracks spread over
 * n_threads host threads; *seconds = wall time of the fits. */
int oracle_fit_run(const ckf_fitting_config* cfg, const ckf_detector* det, const ckf_bfield* field,
                   const ckf_measurements* meas, uint32_t n_tracks, const uint64_t* seed_surface_link,
                   const float* seed_vec, const float* seed_cov, const uint32_t* offsets,
                   const uint32_t* meas_index, int precision, uint32_t n_threads, oracle_fitted* out,
                   double* seconds);
void oracle_fitted_free(oracle_fitted* f);

/* ---- unit-level entry points (FP64) ---- */
double oracle_prob(double chi2, double ndf);
float oracle_prob_f(float chi2, float ndf);

/* Predicted chi2 and gain-matrix update of one bound state with one
 * measurement.  Returns the ckf_fitter_status; out_vec/out_cov filled on
 * success. */
int oracle_kf_update(const double vec[6], const double cov[36], const float meas[2],
                     const float var[2], uint32_t dims, const uint8_t subspace[2],
                     double* chi2, double out_vec[6], double out_cov[36]);

/* Propagate a bound state from surface `start` (index) to the next sensitive
 * surface.  Returns 0 found, 1 exited the detector, 2 aborted. */
int oracle_propagate(const ckf_finding_config* cfg, const ckf_detector* det,
                     const ckf_bfield* field, uint32_t start, const double vec[6],
                     const double cov[36], int apply_start_material, uint32_t* end_surface,
                     double out_vec[6], double out_cov[36], double* path, uint32_t* n_steps);

/* Transport a bound state from surface `start` straight onto surface
 * `target` (either direction along the track; no navigation).  0 on success,
 * 2 aborted. */
int oracle_propagate_direct(const ckf_finding_config* cfg, const ckf_detector* det,
                            const ckf_bfield* field, uint32_t start, uint32_t target,
                            const double vec[6], const double cov[36], double out_vec[6],
                            double out_cov[36]);

/* Greedy ambiguity resolution (ambiguity_oracle.cpp), traccc host
 * greedy_ambiguity_resolution_algorithm restated.  Tracks t own the
 * measurement identifiers meas_ids[offsets[t], offsets[t+1]).  accepted[t]
 * = 1 for every track kept; *n_removed = tracks removed by the iterations.
 * 0 on success. */
int oracle_greedy_resolve(uint32_t n_tracks, const uint32_t* offsets, const uint32_t* meas_ids,
                          const float* pval, uint32_t min_meas_per_track, uint32_t max_iterations,
                          uint32_t max_shared_meas, uint8_t* accepted, uint32_t* n_removed);

/* Seeding (seeding_oracle.cpp), traccc host seeding_algorithm restated in
 * single precision.  Seeds in the reference's order (grid bin, middle
 * spacepoint, then the filter's order). */
typedef struct oracle_seeds {
    uint32_t n_seeds;
    uint32_t n_phi_bins, n_z_bins;
    uint32_t* bottom;
    uint32_t* middle;
    uint32_t* top;
    float* weight;
} oracle_seeds;
void oracle_seedfinder_setup(ckf_seedfinder_config* cfg);
/* 0 ok, 2 = a grid configuration the reference rejects (std::domain_error) */
int oracle_seeding(const ckf_seedfinder_config* finder, const ckf_spacepoint_grid_config* grid,
                   const ckf_seedfilter_config* filter, uint32_t n_sp, const float* global,
                   const float* var_r, const float* var_z, oracle_seeds* out);
void oracle_seeds_free(oracle_seeds* s);
/* Track parameters of one seed (bottom p0, middle p1, top p2) in double
 * precision: vec = (loc0, loc1, phi, theta, q/p, t), cov_diag = the
 * covariance diagonal. */
int oracle_track_params(const ckf_track_params_estimation_config* cfg, const double b[3],
                        const double p0[3], const double p1[3], const double p2[3],
                        const double bottom_local[2], double vec[6], double cov_diag[6]);

/* One RK4 step of a free state [x, y, z, t, Tx, Ty, Tz, q/p] with the
 * transport matrix d(free_out)/d(free_in); for tests. */
void oracle_rk_step(const ckf_bfield* field, double mass, double abs_q, double h,
                    const double free_in[8], double free_out[8], double D[64]);

#ifdef __cplusplus
}
#endif

#endif
