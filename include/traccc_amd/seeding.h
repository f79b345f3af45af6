/*
 * traccc_amd -- spacepoint formation, triplet seeding and track-parameter
 * estimation (C ABI): the steps before the CKF.
 *
 * Replaces:
 *   traccc::host::silicon_pixel_spacepoint_formation_algorithm::operator()(det, measurements)
 *     core/src/seeding/silicon_pixel_spacepoint_formation.hpp:31-58,
 *     core/include/traccc/seeding/impl/spacepoint_formation.ipp:20-48
 *   traccc::host::seeding_algorithm::operator()(spacepoints)
 *     core/include/traccc/seeding/seeding_algorithm.hpp:26-66, core/src/seeding/seeding_algorithm.cpp:12-28
 *     (spacepoint binning core/src/seeding/spacepoint_binning.cpp:17-53 and
 *      core/include/traccc/seeding/spacepoint_binning_helper.hpp:22-127; seed finding
 *      core/src/seeding/seed_finding.cpp:57-121 with doublet_finding.hpp, triplet_finding.hpp,
 *      seed_filtering.cpp and the helpers doublet_finding_helper.hpp,
 *      triplet_finding_helper.hpp, seed_selecting_helper.hpp, detail/triplet_sorter.hpp)
 *   traccc::host::track_params_estimation::operator()(measurements, spacepoints, seeds, bfield)
 *     core/src/seeding/track_params_estimation.cpp:23-90,
 *     core/include/traccc/seeding/track_params_estimation_helper.hpp:47-125
 *
 * Units are traccc's native ones: mm, GeV, e, and the magnetic field in
 * GeV / (e mm) (1 T = 0.000299792458).
 *
 * Same conventions as ckf.h: plain C, ckf_status returns, thread-local
 * message in ckf_last_error().
 */
#ifndef TRACCC_AMD_SEEDING_H
#define TRACCC_AMD_SEEDING_H

#include <stdint.h>

#include "traccc_amd/ckf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of traccc::seedfinder_config
 * (core/include/traccc/seeding/detail/seeding_config.hpp:17-141).  The last
 * six fields are derived: ckf_seedfinder_config_setup() computes them from
 * the others as the reference's setup() does; like the reference, the
 * library uses them as they are (changing minPt afterwards without setup
 * leaves them at their old values). */
typedef struct ckf_seedfinder_config {
    float zMin, zMax, rMax, rMin;                 /* -2000, 2000, 200, 33 mm */
    float collisionRegionMin, collisionRegionMax; /* -250, 250 mm */
    float phiMin, phiMax;                         /* -pi, pi */
    float minPt;                                  /* 0.5 GeV */
    float cotThetaMax;                            /* 27.2845 */
    float deltaRMin, deltaRMax;                   /* 20, 80 mm */
    float deltaZMax;                              /* 450 mm */
    float impactMax;                              /* 10 mm */
    float sigmaScattering;                        /* 3 */
    float maxPtScattering;                        /* 10 GeV */
    uint32_t maxSeedsPerSpM;                      /* 5 */
    float bFieldInZ;                              /* 1.99724 T */
    float beamPos[2];                             /* 0, 0 */
    float radLengthPerSeed;                       /* 0.05 */
    float zAlign, rAlign;                         /* 0 (not used by the algorithm) */
    float sigmaError;                             /* 5 (not used by the algorithm) */
    int32_t phiBinDeflectionCoverage;             /* 1 */
    uint32_t neighbor_scope[2];                   /* 1, 1 bins on either side */
    /* derived (setup) */
    float highland, maxScatteringAngle2, pTPerHelixRadius, minHelixDiameter2, minHelixRadius,
        pT2perRadius;
} ckf_seedfinder_config;

/* Mirror of traccc::spacepoint_grid_config (seeding_config.hpp:143-200):
 * the reference constructs it from a finder config (copying these fields),
 * after which either can change independently. */
typedef struct ckf_spacepoint_grid_config {
    float bFieldInZ, minPt, rMax, zMax, zMin, deltaRMax, cotThetaMax, impactMax, phiMin, phiMax;
    int32_t phiBinDeflectionCoverage;
} ckf_spacepoint_grid_config;

/* Mirror of traccc::seedfilter_config (seeding_config.hpp:202-238). */
typedef struct ckf_seedfilter_config {
    float deltaInvHelixDiameter;    /* 0.00003 / mm */
    float impactWeightFactor;       /* 1 */
    float compatSeedWeight;         /* 200 */
    float deltaRMin;                /* 5 mm */
    uint32_t compatSeedLimit;       /* 2 */
    float good_spB_min_radius;      /* 150 mm */
    float good_spB_weight_increase; /* 400 */
    float good_spT_max_radius;      /* 150 mm */
    float good_spT_weight_increase; /* 200 */
    float good_spB_min_weight;      /* 380 */
    float seed_min_weight;          /* 200 */
    float spB_min_radius;           /* 43 mm */
} ckf_seedfilter_config;

/* Mirror of traccc::track_params_estimation_config
 * (core/include/traccc/seeding/detail/track_params_estimation_config.hpp:17-32). */
typedef struct ckf_track_params_estimation_config {
    float initial_sigma[6];      /* 1 mm, 1 mm, 1 deg, 1 deg, 0 e/GeV, 1 ns */
    float initial_sigma_qopt;    /* 0.1 e/GeV */
    float initial_sigma_pt_rel;  /* 0.1 */
    float initial_inflation[6];  /* 1, 1, 1, 1, 1, 100 */
} ckf_track_params_estimation_config;

/* traccc edm::spacepoint_collection (core/include/traccc/edm/spacepoint_collection.hpp):
 * global position, variances, and the measurement(s) it was made of
 * (measurement_index_2 = 0xffffffff for a pixel spacepoint). */
typedef struct ckf_spacepoints {
    uint32_t size;
    const float* global;                 /* [size][3] */
    const float* radius_variance;        /* [size] (NULL: 0) */
    const float* z_variance;             /* [size] (NULL: 0) */
    const uint32_t* measurement_index_1; /* [size] (NULL: none) */
} ckf_spacepoints;

/* traccc edm::seed_collection (core/include/traccc/edm/seed_collection.hpp):
 * spacepoint indices of the bottom, middle and top spacepoint, and the
 * seed's weight.  Caller-owned, sized from ckf_seeding_result_size(). */
typedef struct ckf_seeds {
    uint32_t size;
    uint32_t* bottom;
    uint32_t* middle;
    uint32_t* top;
    float* weight;
} ckf_seeds;

void ckf_seedfinder_config_default(ckf_seedfinder_config* cfg);
/* the reference's seedfinder_config::setup(): the derived fields */
void ckf_seedfinder_config_setup(ckf_seedfinder_config* cfg);
/* the reference's spacepoint_grid_config(const seedfinder_config&) */
void ckf_spacepoint_grid_config_from(const ckf_seedfinder_config* finder, ckf_spacepoint_grid_config* out);
void ckf_seedfilter_config_default(ckf_seedfilter_config* cfg);
void ckf_track_params_estimation_config_default(ckf_track_params_estimation_config* cfg);

typedef struct ckf_seeding ckf_seeding;

/* traccc: seeding_algorithm(finder_config, grid_config, filter_config, mr).
 * A grid configuration the reference rejects with std::domain_error (the
 * minimum helix radius below rMax / 2, or no positive phi bin width) gives
 * CKF_ERR_INVALID_ARGUMENT. */
ckf_status ckf_seeding_create(const ckf_seedfinder_config* finder, const ckf_spacepoint_grid_config* grid,
                              const ckf_seedfilter_config* filter, ckf_seeding** out);
ckf_status ckf_seeding_destroy(ckf_seeding* s);
ckf_status ckf_seeding_set_stream(ckf_seeding* s, void* hip_stream);

/* The operator on host spacepoints (copied to HBM); the seeds stay in the
 * object until the next call: *n_seeds is their number. */
ckf_status ckf_seeding_run(ckf_seeding* s, const ckf_spacepoints* spacepoints, uint32_t* n_seeds);
/* The same on spacepoints already in HBM (device pointers). */
ckf_status ckf_seeding_run_device(ckf_seeding* s, const ckf_spacepoints* spacepoints, uint32_t* n_seeds);
ckf_status ckf_seeding_result_size(const ckf_seeding* s, uint32_t* n_seeds);
/* Copy the last call's seeds to caller host arrays (any may be NULL). */
ckf_status ckf_seeding_result_download(const ckf_seeding* s, ckf_seeds* out);

/* Spacepoint formation on the device: one spacepoint per 2D measurement of
 * an event of `ev` (measurement order), global = surface centre + loc0 u +
 * loc1 v, variances 0.  global: caller host array [n][3] or NULL;
 * measurement_index: [n] or NULL; *n = spacepoints made.  Call with both
 * NULL to get *n first. */
ckf_status ckf_spacepoint_formation(const ckf_device_detector* det, const ckf_device_event* ev,
                                    uint32_t event, float* global, uint32_t* measurement_index,
                                    uint32_t* n);

/* Track parameters at each seed's bottom spacepoint (host arrays):
 * seeds (bottom/middle/top indices into the spacepoints), the spacepoints'
 * global positions, each seed's bottom measurement (local position and
 * surface link, from measurement_index_1) and the field vector.  Output
 * traccc bound parameters (loc0, loc1, phi, theta, q/p, time) with the
 * reference's diagonal covariance. */
ckf_status ckf_estimate_track_params(const ckf_track_params_estimation_config* cfg,
                                     const ckf_spacepoints* spacepoints, const ckf_seeds* seeds,
                                     const ckf_measurements* measurements, const float bfield[3],
                                     uint64_t* surface_link, float* vector, float* covariance);

#ifdef __cplusplus
}
#endif

#endif /* TRACCC_AMD_SEEDING_H */
