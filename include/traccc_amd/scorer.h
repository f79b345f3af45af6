/*
 * traccc_amd -- learned branch scorer for the CKF (C ABI).
 *
 * No counterpart in traccc: a small MLP (12 -> 64 -> 64 -> 1, GELU, BF16
 * weights, FP32 accumulation) that scores a CKF branch candidate -- one
 * (branch, measurement) pair that passed the chi2 cut -- by its chance to end
 * in a found track.  It is trained on survive / die labels of branches logged
 * by the plain CKF itself (ckf_find_tracks_logged) with the project's own
 * BF16-MFMA forward / backward kernels, and prunes candidates before the
 * Kalman update (ckf_algorithm_set_scorer).  The plain CKF stays available
 * and is the yardstick of finding efficiency.
 *
 * Features (in this order, FP32, normalised in the kernel with the scorer's
 * mean / std):
 *   0 pull in loc0   1 pull in loc1 (0 for 1-D)   2 predicted chi2
 *   3 layer index / layer count   4 q/p (1/GeV)   5 theta
 *   6 holes so far   7 branch depth (CKF step)   8 chi2 sum / ndf sum so far
 *   9 standing: the branch's running chi2 / ndf with this candidate minus
 *     the lowest such value over the rank-0 candidates of all branches of
 *     the seed at this step (0 for the seed's best branch)
 *   10 rank of the candidate among the branch's candidates by chi2
 *   11 log(1 + candidates passing the chi2 cut)
 *
 * Parameters (flat FP32, ckf_scorer_n_params() = 5057):
 *   W1[12][64] b1[64] W2[64][64] b2[64] W3[64] b3[1]
 *   (W[i][j]: input i -> output j).
 */
#ifndef TRACCC_AMD_SCORER_H
#define TRACCC_AMD_SCORER_H

#include <stdint.h>

#include "traccc_amd/ckf.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CKF_SCORER_FEATURES 12u
#define CKF_SCORER_HIDDEN 64u

typedef struct ckf_scorer ckf_scorer;

typedef struct ckf_scorer_train_config {
    uint32_t epochs;        /* passes over the samples                       */
    uint32_t batch;         /* samples per optimiser step (multiple of 64)   */
    float learning_rate;    /* Adam                                          */
    float beta1, beta2, eps;
    float weight_decay;     /* decoupled (AdamW)                             */
    float pos_weight;       /* weight of survive labels in the loss          */
    uint64_t shuffle_seed;  /* sample order per epoch                        */
} ckf_scorer_train_config;

typedef struct ckf_scorer_train_stats {
    uint32_t steps;         /* optimiser steps taken                         */
    float first_loss;       /* mean loss of the first epoch                  */
    float last_loss;        /* mean loss of the last epoch                   */
    double seconds;         /* device time of the training kernels           */
    double tflops;          /* achieved (algorithmic) TFLOP/s of fwd + bwd    */
} ckf_scorer_train_stats;

uint32_t ckf_scorer_n_params(void);
void ckf_scorer_train_config_default(ckf_scorer_train_config* cfg);

/* Random initialisation (scaled uniform, seeded), identity normalisation. */
ckf_status ckf_scorer_create(uint64_t seed, ckf_scorer** out);
ckf_status ckf_scorer_destroy(ckf_scorer* s);

/* Host copies of the FP32 master parameters and the feature normalisation
 * (mean[12], std[12]); either pointer may be NULL. */
ckf_status ckf_scorer_set_params(ckf_scorer* s, const float* params, const float* mean,
                                 const float* std_dev);
ckf_status ckf_scorer_get_params(const ckf_scorer* s, float* params, float* mean, float* std_dev);

/* Device arrays: features [n][12] FP32, labels [n] u8 (1 = survives), scores
 * [n] FP32 (logits).  Calls are ordered on the stream (hipStream_t as void*,
 * NULL = default) and return after it has finished. */
ckf_status ckf_scorer_fit_normalisation(ckf_scorer* s, const float* d_features, uint32_t n,
                                        void* stream);
ckf_status ckf_scorer_forward(const ckf_scorer* s, const float* d_features, uint32_t n,
                              float* d_scores, void* stream);
/* Gradient of the mean (weighted) binary cross-entropy over the n samples with
 * respect to the parameters, into the host array grad[n_params]; *loss gets
 * the mean loss.  For gradient checks; no parameter changes. */
ckf_status ckf_scorer_gradient(const ckf_scorer* s, const float* d_features,
                               const uint8_t* d_labels, uint32_t n, float pos_weight,
                               float* grad, float* loss, void* stream);
ckf_status ckf_scorer_train(ckf_scorer* s, const float* d_features, const uint8_t* d_labels,
                            uint32_t n, const ckf_scorer_train_config* cfg,
                            ckf_scorer_train_stats* stats, void* stream);

/* ckf_scorer_gradient without any device allocation -- no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, training loops that forbid hipMalloc).  Same gradient and loss,
 * bit for bit, into caller DEVICE memory: d_grad [ckf_scorer_n_params()]
 * floats, d_loss [1] float.  The per-tile partial gradients go to a fixed
 * device array of the library (20.7 MB per GPU, held until the process
 * ends), so n is limited to CKF_SCORER_NOALLOC_MAX_SAMPLES
 * (CKF_ERR_INVALID_ARGUMENT above it).  Ordered on the stream, no host
 * synchronisation.  Every call on one GPU shares that array, and each call
 * resets the scorer's training state: calls that overlap on one GPU (or
 * overlap a ckf_scorer_train of the same scorer) race. */
#define CKF_SCORER_NOALLOC_MAX_SAMPLES 65536u
ckf_status ckf_scorer_gradient_noalloc(const ckf_scorer* s, const float* d_features,
                                       const uint8_t* d_labels, uint32_t n, float pos_weight,
                                       float* d_grad, float* d_loss, void* stream);

/* ---- in the CKF ----------------------------------------------------------- */
/* Prune, before the Kalman update, every candidate whose logit is below
 * `threshold`: each input branch lists its max_num_branches_per_surface + 1
 * best candidates by (chi2, measurement index) -- the plain CKF's selection
 * and one more in place of a failed update -- the scorer scores the list,
 * and the update walks the surviving entries in order as the plain CKF
 * walks all candidates.  scorer == NULL restores the plain CKF.  The scorer must
 * outlive its use by the algorithm. */
ckf_status ckf_algorithm_set_scorer(ckf_algorithm* alg, const ckf_scorer* scorer, float threshold);

/* The plain CKF with its branches logged (training data of the scorer):
 * every measurement link of the run, in link order, gives one sample --
 * its 12 features at creation into d_features[n][12] and, into d_labels[n],
 * 1 when the link lies on its seed's best found track candidate (most
 * measurements, then the lowest chi2 sum), 0 otherwise: the branch that the
 * seed's best track needs survives, every other branch dies.  Writes at most
 * `capacity` rows; *n_samples gets the number of samples of the run. */
ckf_status ckf_find_tracks_logged(ckf_algorithm* alg, const ckf_device_detector* det,
                                  const ckf_device_event* ev, float* d_features,
                                  uint8_t* d_labels, uint32_t capacity, uint32_t* n_samples,
                                  ckf_device_tracks** out);

#ifdef __cplusplus
}
#endif

#endif /* TRACCC_AMD_SCORER_H */
