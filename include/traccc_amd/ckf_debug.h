/*
 * Diagnostic entry points of the traccc_amd CKF library.
 *
 * No counterpart in the reference: these run single pieces of the device
 * step pipeline on caller arrays so tests can compare them with a CPU oracle
 * (the Kalman update of traccc's gain_matrix_updater,
 * core/include/traccc/fitting/kalman_filter/gain_matrix_updater.hpp:69-181,
 * and the propagation of details/combinatorial_kalman_filter.hpp:528-627).
 * All pointers are host pointers; each call uploads, runs one kernel and
 * downloads.
 */
#ifndef TRACCC_AMD_CKF_DEBUG_H
#define TRACCC_AMD_CKF_DEBUG_H

#include <stdint.h>

#include "traccc_amd/ckf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* For n (state, measurement) pairs: predicted chi2 and gain-matrix update.
 * vec [n][6], cov [n][36] (row-major, symmetric), meas/var [n][2], dims [n],
 * subspace [n][2].  Outputs: chi2 [n], status [n] (ckf_fitter_status),
 * out_vec [n][6], out_cov [n][36]. */
ckf_status ckf_debug_kf_update(uint32_t n, const float* vec, const float* cov, const float* meas,
                               const float* var, const uint32_t* dims, const uint8_t* subspace,
                               float* chi2, int32_t* status, float* out_vec, float* out_cov);

/* Propagate n bound states from surface index start[i] to the next sensitive
 * surface with the CKF's propagation settings (material of the start surface
 * applied first when apply_start_material != 0, of the end surface always).
 * status[i]: 0 found, 1 left the detector, 2 aborted. */
ckf_status ckf_debug_propagate(const ckf_finding_config* cfg, const ckf_detector* det,
                               const ckf_bfield* field, uint32_t n, const uint32_t* start,
                               const float* vec, const float* cov, int apply_start_material,
                               int32_t* status, uint32_t* end_surface, float* out_vec,
                               float* out_cov, float* path, uint32_t* n_steps);

/* Per-kernel device time of the last ckf_find_tracks call, measured with HIP
 * events on the algorithm's stream around each kernel group of every step
 * (find; duplicate removal; propagation; step advance).  While enabled the
 * step loop is launched step by step instead of as a captured graph. */
ckf_status ckf_debug_set_kernel_timing(ckf_algorithm* alg, int enable);

/* Propagation schedule: lanes that navigate one link together in the
 * propagation kernel (1, 2, 4, 8 or 16; 0 = autotuned per seed-count class, the
 * default).  *lanes receives the schedule of the last call. */
ckf_status ckf_debug_set_schedule(ckf_algorithm* alg, uint32_t lanes_per_link);
ckf_status ckf_debug_get_schedule(const ckf_algorithm* alg, uint32_t* lanes_per_link);

/* One 16x16x16 BF16 matrix-core product D = A B (row-major FP32 host arrays,
 * inputs rounded to BF16) through the fragment layout the scorer kernels
 * use; checks that layout on the device. */
ckf_status ckf_debug_mfma_layout(const float* A, const float* B, float* D);
ckf_status ckf_debug_kernel_times(const ckf_algorithm* alg, double* ms_find, double* ms_dedup,
                                  double* ms_propagate, double* ms_other, uint32_t* n_steps);
/* The find group of the last timed call split into its two kernels: the
 * find kernel (plain, or scorer-pruned: candidate lists, the MLP and the
 * updates in one kernel) and its link compaction; their sum is ms_find. */
ckf_status ckf_debug_find_split(const ckf_algorithm* alg, double* ms_find_kernel,
                                double* ms_find_emit);
/* The propagation group of the last timed call split into the propagation
 * kernel (k_propagate at the schedule's team size) and its compaction; their
 * sum is ms_propagate. */
ckf_status ckf_debug_propagate_split(const ckf_algorithm* alg, double* ms_propagate_kernel,
                                     double* ms_propagate_emit);

/* Candidate-verdict log of scorer-pruned calls, for replay by a CPU
 * checker: while the capacity is above 0 every ckf_find_tracks call with a
 * scorer records one 64-bit key per scored candidate, in no particular
 * order.  The key's low 63 bits are splitmix64 chained over the seed index,
 * the step, the measurement index of every earlier link of the branch
 * (latest first, holes as 0xffffffff) and the candidate's measurement index
 * (h = mix(h ^ v), from h = 0); bit 63 is set when the candidate was kept.
 * The log restarts with each call (and with each internal rerun after a
 * buffer regrow).  ckf_debug_get_verdicts copies up to `capacity` keys of the
 * last call and gives the number recorded (which may exceed the log's
 * capacity: then enlarge it and call again). */
ckf_status ckf_debug_set_verdict_log(ckf_algorithm* alg, uint32_t capacity);
ckf_status ckf_debug_get_verdicts(const ckf_algorithm* alg, uint64_t* keys, uint32_t capacity,
                                  uint32_t* n_records);

#ifdef __cplusplus
}
#endif

#endif
