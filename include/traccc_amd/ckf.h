/*
 * traccc_amd -- MI355X-native combinatorial Kalman filter (CKF) track finding.
 *
 * C ABI of the drop-in boundary.  Every entry point here replaces one piece of
 * traccc's track-finding operator:
 *
 *   traccc::host::combinatorial_kalman_filter_algorithm
 *     core/include/traccc/finding/combinatorial_kalman_filter_algorithm.hpp:34-72
 *   traccc::device::combinatorial_kalman_filter_algorithm
 *     device/common/include/traccc/finding/device/combinatorial_kalman_filter_algorithm.hpp:49-90
 *
 * Plain C types only: pointers, sizes, PODs.  No C++ exception crosses this
 * boundary; every call returns a ckf_status and leaves a thread-local message
 * readable through ckf_last_error().
 *
 * Units follow traccc/detray: mm, GeV, ns, elementary charge; magnetic field
 * values handed in through ckf_bfield are in tesla.
 */
#ifndef TRACCC_AMD_CKF_H
#define TRACCC_AMD_CKF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CKF_ABI_VERSION 1u

/* ------------------------------------------------------------------------- */
/* Status codes (traccc throws std::invalid_argument / runtime errors; the C  */
/* ABI turns those into codes).                                               */
/* ------------------------------------------------------------------------- */
typedef enum ckf_status {
    CKF_OK = 0,
    CKF_ERR_INVALID_ARGUMENT = 1, /* traccc: std::invalid_argument from ctor */
    CKF_ERR_UNSUPPORTED = 2,      /* traccc: unsupported detector/field type */
    CKF_ERR_DEVICE = 3,           /* HIP runtime error */
    CKF_ERR_OUT_OF_MEMORY = 4,
    CKF_ERR_PRECONDITION = 5,     /* measurements not grouped by surface etc. */
    CKF_ERR_INTERNAL = 6
} ckf_status;

/* traccc smoother_type (core/include/traccc/finding/finding_config.hpp:26-31) */
typedef enum ckf_smoother_type {
    CKF_SMOOTHER_NONE = 0,
    CKF_SMOOTHER_MBF = 1,
    CKF_SMOOTHER_KALMAN = 2,
    CKF_SMOOTHER_UNKNOWN = 3
} ckf_smoother_type;

/* traccc kalman_fitter_status (core/include/traccc/fitting/status_codes.hpp:14-29) */
typedef enum ckf_fitter_status {
    CKF_FIT_SUCCESS = 0,
    CKF_FIT_ERROR_QOP_ZERO = 1,
    CKF_FIT_ERROR_THETA_POLE = 2,
    CKF_FIT_ERROR_INVERSION = 3,
    CKF_FIT_ERROR_UPDATER_INVALID_COVARIANCE = 8,
    CKF_FIT_ERROR_OTHER = 12
} ckf_fitter_status;

/* Propagation settings, named after detray::propagation::config's stepping and
 * navigation members (examples/options/src/track_propagation.cpp:23-156). */
typedef struct ckf_propagation_config {
    /* stepping */
    float step_constraint;   /* mm, hard cap on one RK step                   */
    float rk_error_tol;      /* mm, RK4 local error target per step           */
    float min_stepsize;      /* mm, smallest RK step taken on approach        */
    float path_limit;        /* mm, total path allowed for one propagation    */
    /* navigation */
    float overstep_tolerance;   /* mm (<= 0): a surface this far behind still counts as ahead */
    float min_mask_tolerance;   /* mm, candidate pre-selection margin at zero distance */
    float max_mask_tolerance;   /* mm, cap of the pre-selection margin         */
    float mask_tolerance_scalor;/* margin growth per mm of straight-line distance */
    float on_surface_tolerance; /* mm, plane distance below which the track sits on the surface */
    uint32_t search_window[2];  /* neighbour bins scanned on each side (phi, long) */
} ckf_propagation_config;

/* Mirror of traccc::finding_config (core/include/traccc/finding/finding_config.hpp:70-148),
 * field for field, same names and defaults (see ckf_finding_config_default). */
typedef struct ckf_finding_config {
    uint32_t max_num_branches_per_seed;        /* 10   */
    uint32_t max_num_branches_per_surface;     /* 1    */
    uint32_t min_track_candidates_per_track;   /* 3    */
    uint32_t max_track_candidates_per_track;   /* 100  */
    uint32_t max_num_skipping_per_cand;        /* 3    */
    uint32_t max_num_consecutive_skipped;      /* 1    */
    uint32_t max_num_tracks_per_measurement;   /* 0    */
    float min_measurement_voting_fraction;     /* 0.5  */
    uint32_t run_pkf;                          /* 0    */
    uint32_t run_smoother;                     /* ckf_smoother_type, CKF_SMOOTHER_MBF */
    float min_step_length_for_next_surface;    /* 1.2 mm */
    uint32_t max_step_counts_for_next_surface; /* 100  */
    float chi2_max;                            /* 30   */
    ckf_propagation_config propagation;
    float min_p;                               /* 0.1 GeV */
    float min_pT;                              /* 0.6 GeV */
    int32_t ptc_pdg;                           /* 211 (pi+) */
    float ptc_mass;                            /* GeV */
    float ptc_charge;                          /* e   */
    uint32_t duplicate_removal_minimum_length; /* 5    */
    uint32_t initial_links_per_seed;           /* 100  */
} ckf_finding_config;

/* ------------------------------------------------------------------------- */
/* Detector description.  Replaces traccc::host_detector / detector_buffer    */
/* (core/include/traccc/geometry/host_detector.hpp:21-86,                    */
/*  core/include/traccc/geometry/detector_buffer.hpp:21-55) for the toy       */
/* cylindrical geometry: planar modules grouped in navigation layers.        */
/* ------------------------------------------------------------------------- */
typedef enum ckf_bound_kind { CKF_BOUND_RECTANGLE = 0, CKF_BOUND_TRAPEZOID = 1 } ckf_bound_kind;
typedef enum ckf_layer_kind {
    CKF_LAYER_BARREL = 0, /* modules on a cylinder, (phi, z) grid               */
    CKF_LAYER_RING = 1,   /* modules on a disk ring, phi grid                   */
    CKF_LAYER_PLANE = 2   /* a slab along `normal`; every module is a candidate */
} ckf_layer_kind;

typedef struct ckf_surface {
    float center[3];
    float axis_u[3];     /* local loc0 axis                                  */
    float axis_v[3];     /* local loc1 axis                                  */
    float normal[3];     /* axis_u x axis_v                                  */
    float bound[3];      /* rectangle: half_u, half_v, 0                     */
                         /* trapezoid: half_u at v=-half_v, at v=+half_v, half_v */
    float thickness_x0;  /* x/X0 at normal incidence; 0 = no material        */
    uint32_t bound_kind; /* ckf_bound_kind                                   */
    uint32_t layer;      /* index into ckf_detector.layers                   */
    uint32_t sensitive;  /* 1 if measurements may live on it                 */
    uint32_t reserved;
    uint64_t identifier; /* geometry identifier; strictly increasing with the index */
} ckf_surface;

typedef struct ckf_layer {
    uint32_t kind;          /* ckf_layer_kind                                  */
    uint32_t first_surface; /* module (iphi, il) = first + iphi * n_long + il   */
    uint32_t n_phi;
    uint32_t n_long;        /* barrel: modules along z; ring: 1                 */
    float phi0;             /* phi of the iphi = 0 module centre               */
    float long0;            /* barrel: z of the il = 0 module centre           */
    float long_pitch;       /* barrel: z pitch of module centres               */
    float env_r[2];         /* radial extent of the module material; plane: slab range along normal */
    float env_z[2];         /* longitudinal extent of the module material      */
    float normal[3];        /* plane layers only                               */
} ckf_layer;

typedef struct ckf_detector {
    const ckf_surface* surfaces;
    uint32_t n_surfaces;
    const ckf_layer* layers;
    uint32_t n_layers;
} ckf_detector;

/* Magnetic field.  Replaces traccc::magnetic_field (core/include/traccc/bfield/magnetic_field.hpp:21-115)
 * with its two backends: constant, and an (r, z) map with bilinear lookup. */
typedef enum ckf_bfield_kind { CKF_BFIELD_CONSTANT = 0, CKF_BFIELD_RZ_MAP = 1 } ckf_bfield_kind;

typedef struct ckf_bfield {
    uint32_t kind;       /* ckf_bfield_kind                                   */
    float b[3];          /* constant field, tesla                             */
    const float* map;    /* rz map: [n_z][n_r][2] = (B_r, B_z) in tesla       */
    uint32_t n_r, n_z;
    float r_min, z_min;  /* grid origin, mm                                    */
    float dr, dz;        /* grid pitch, mm                                     */
} ckf_bfield;

/* ------------------------------------------------------------------------- */
/* Event data model.                                                          */
/* ------------------------------------------------------------------------- */
/* traccc::edm::measurement_collection columns
 * (core/include/traccc/edm/measurement_collection.hpp:241-260).  Rows must be
 * grouped (sorted) by surface_link, as traccc requires
 * (core/include/traccc/finding/details/combinatorial_kalman_filter.hpp:94-95). */
typedef struct ckf_measurements {
    uint32_t size;
    const float* local_position;  /* [size][2] */
    const float* local_variance;  /* [size][2] */
    const uint32_t* dimensions;   /* 1 or 2    */
    const float* time;            /* optional (may be NULL) */
    const float* diameter;        /* optional */
    const uint32_t* identifier;   /* optional */
    const uint64_t* surface_link; /* geometry identifier of the surface */
    const uint8_t* subspace;      /* [size][2] bound indices measured  */
    const uint32_t* cluster_index;/* optional */
} ckf_measurements;

/* traccc bound_track_parameters collection (core/include/traccc/edm/track_parameters.hpp:28-29).
 * Parameter order: loc0, loc1, phi, theta, q/p, time. */
typedef struct ckf_bound_params {
    uint32_t size;
    const uint64_t* surface_link;
    const float* vector;      /* [size][6]      */
    const float* covariance;  /* [size][6][6] row-major */
} ckf_bound_params;

/* traccc track_container output (core/include/traccc/edm/track_container.hpp:22-152,
 * track_collection.hpp:183-198, track_constituent_link.hpp:20-53).  Caller-owned
 * host buffers sized from ckf_tracks_size(). */
typedef struct ckf_track_candidates {
    uint32_t n_tracks;
    uint32_t n_constituents;
    uint16_t* fit_outcome;          /* [n_tracks]; traccc track_fit_outcome::UNKNOWN = 0 */
    uint64_t* params_surface_link;  /* [n_tracks] */
    float* params_vector;           /* [n_tracks][6]  (the seed, as traccc) */
    float* params_covariance;       /* [n_tracks][36] */
    float* ndf;                     /* [n_tracks] */
    float* chi2;                    /* [n_tracks] */
    float* pval;                    /* [n_tracks] */
    uint32_t* nholes;               /* [n_tracks] */
    uint32_t* constituent_offsets;  /* [n_tracks + 1] */
    uint16_t* constituent_type;     /* [n_constituents]; 0 = measurement */
    uint32_t* constituent_index;    /* [n_constituents]; index into the caller's measurements */
} ckf_track_candidates;

/* Opaque handles */
typedef struct ckf_algorithm ckf_algorithm;             /* one configured CKF        */
typedef struct ckf_device_detector ckf_device_detector; /* geometry + field in HBM   */
typedef struct ckf_device_event ckf_device_event;       /* one event's inputs in HBM */
typedef struct ckf_device_tracks ckf_device_tracks;     /* track candidates in HBM   */

/* Run statistics of the last call on an algorithm (diagnostics and bench). */
typedef struct ckf_run_stats {
    uint32_t n_steps;        /* CKF steps executed                             */
    uint64_t n_links;        /* links created over all steps                   */
    uint64_t n_propagations; /* propagate_to_next_surface calls                */
    uint64_t n_tips;         /* tips before the length filter                  */
    uint32_t n_regrows;      /* buffer regrow reruns (traccc: link buffer resize) */
    uint32_t reserved;
    uint64_t n_scored;       /* candidates scored by the branch scorer (0 without) */
} ckf_run_stats;

/* ---- library ------------------------------------------------------------ */
uint32_t ckf_abi_version(void);
const char* ckf_last_error(void);
void ckf_finding_config_default(ckf_finding_config* cfg);

/* ---- algorithm object (traccc ctor: finding_config + memory resource) ----- */
/* Replaces combinatorial_kalman_filter_algorithm(config, mr, logger):
 * core/src/finding/combinatorial_kalman_filter_algorithm.cpp:21-32 and the
 * device ctor's checks device/common/src/finding/combinatorial_kalman_filter_algorithm.cpp:33-46.
 * Returns CKF_ERR_INVALID_ARGUMENT where traccc throws std::invalid_argument. */
ckf_status ckf_algorithm_create(const ckf_finding_config* cfg, ckf_algorithm** out);
ckf_status ckf_algorithm_destroy(ckf_algorithm* alg);
/* HIP stream (hipStream_t passed as void*) every launch of this object uses;
 * NULL = the default stream.  traccc: cuda::stream_wrapper in the CUDA ctor
 * (device/cuda/src/finding/combinatorial_kalman_filter_algorithm.cpp:13-20). */
ckf_status ckf_algorithm_set_stream(ckf_algorithm* alg, void* hip_stream);
ckf_status ckf_algorithm_last_stats(const ckf_algorithm* alg, ckf_run_stats* stats);

/* ---- device-resident inputs ---------------------------------------------- */
/* traccc::buffer_from_host_detector (core/include/traccc/geometry/detector_buffer.hpp:96)
 * plus the device magnetic field (device/cuda/src/utils/magnetic_field_types.hpp:24-51). */
ckf_status ckf_device_detector_create(const ckf_detector* det, const ckf_bfield* field,
                                      ckf_device_detector** out);
ckf_status ckf_device_detector_destroy(ckf_device_detector* det);

/* Copy events' measurements and seeds to HBM (the vecmem::copy setup of the
 * traccc device chain).  n_events >= 1 independent events form one batch that
 * a single ckf_find_tracks call processes in one step pipeline.  Checks the
 * grouped-by-surface precondition (traccc device: input_is_valid,
 * device/cuda/src/finding/combinatorial_kalman_filter_algorithm.cu:39-47) and
 * maps surface identifiers to detector surfaces. */
ckf_status ckf_device_event_create(const ckf_device_detector* det,
                                   const ckf_measurements* meas,
                                   const ckf_bound_params* seeds, uint32_t n_events,
                                   ckf_device_event** out);
ckf_status ckf_device_event_destroy(ckf_device_event* ev);

/* ---- the operator --------------------------------------------------------- */
/* traccc: combinatorial_kalman_filter_algorithm::operator()(det, field, measurements, seeds)
 * core/src/finding/combinatorial_kalman_filter_algorithm.cpp:35-59 (host) and
 * device/common/src/finding/combinatorial_kalman_filter_algorithm.cpp:52-704 (device).
 * Inputs and outputs stay in HBM; the call is ordered on the algorithm's stream
 * and returns once the result sizes are known.  For a batch, the tracks of
 * event e are [track_offsets[e], track_offsets[e+1]) (ckf_tracks_event_offsets). */
ckf_status ckf_find_tracks(ckf_algorithm* alg, const ckf_device_detector* det,
                           const ckf_device_event* ev, ckf_device_tracks** out);

/* Host-pointer convenience form of the operator for one event (uploads,
 * runs; download with ckf_tracks_download). */
ckf_status ckf_find_tracks_host(ckf_algorithm* alg, const ckf_detector* det,
                                const ckf_bfield* field, const ckf_measurements* meas,
                                const ckf_bound_params* seeds, ckf_device_tracks** out);

/* ---- results ---------------------------------------------------------------- */
ckf_status ckf_tracks_size(const ckf_device_tracks* trk, uint32_t* n_tracks,
                           uint32_t* n_constituents);
/* Per-event split of a batched result: n_events + 1 entries each (either may be NULL). */
ckf_status ckf_tracks_event_offsets(const ckf_device_tracks* trk, uint32_t* n_events,
                                    uint32_t* track_offsets, uint32_t* constituent_offsets);
/* Copy into caller-owned host buffers sized from ckf_tracks_size(). */
ckf_status ckf_tracks_download(const ckf_device_tracks* trk, ckf_track_candidates* out);
ckf_status ckf_tracks_destroy(ckf_device_tracks* trk);

/* Device blocks of destroyed results are kept (up to 4 GiB) for the next
 * calls' results; this frees them.  No counterpart in the reference (traccc's
 * results are owned by the caller's vecmem memory resource). */
ckf_status ckf_release_cached_memory(void);

#ifdef __cplusplus
}
#endif

#endif /* TRACCC_AMD_CKF_H */
