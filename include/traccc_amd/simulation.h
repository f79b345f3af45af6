/*
 * traccc_amd synthetic-input toolkit (C ABI).
 *
 * Not part of the drop-in boundary: this builds the toy detectors, the
 * solenoid field map and seeded particle-gun events that the CKF consumes,
 * the role traccc's simulation/ and tests/common fixtures play
 * (simulation/include/traccc/simulation/simulator.hpp:25-170,
 *  simulation/include/traccc/simulation/measurement_smearer.hpp:26-124,
 *  performance/include/traccc/utils/event_data.hpp:79-118).
 * Library-owned arrays are released with the matching *_free call.
 */
#ifndef TRACCC_AMD_SIMULATION_H
#define TRACCC_AMD_SIMULATION_H

#include <stdint.h>

#include "traccc_amd/ckf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tamd_detector_preset {
    TAMD_DET_BARREL4 = 0,        /* 4 barrel layers r = 32/72/116/172 mm, |z| < 500 mm   */
    TAMD_DET_BARREL4_DISK4 = 1,  /* + 2 x 4 endcap disks at |z| = 600..1500 mm           */
    TAMD_DET_BARREL5_DISK6 = 2,  /* 5 barrel layers + 2 x 6 disks at |z| = 600..1500 mm  */
    TAMD_DET_TELESCOPE = 3       /* planes normal to x at x = (k + 1) * spacing          */
} tamd_detector_preset;

typedef struct tamd_detector {
    ckf_surface* surfaces;
    uint32_t n_surfaces;
    ckf_layer* layers;
    uint32_t n_layers;
} tamd_detector;

/* thickness_x0: material per module (x/X0); telescope_planes / spacing only
 * used by TAMD_DET_TELESCOPE. */
int tamd_detector_build(uint32_t preset, float thickness_x0, uint32_t telescope_planes,
                        float telescope_spacing, tamd_detector* out);
void tamd_detector_free(tamd_detector* det);

/* Ideal finite solenoid (current sheet of radius `radius`, full length
 * `length`) scaled to `b_centre` tesla at the origin, tabulated on an (r, z)
 * grid: r in [0, r_max], z in [-z_max, z_max], pitch `pitch` (mm).
 * Writes the ckf_bfield description; the table is library-owned. */
int tamd_solenoid_map_build(float radius, float length, float b_centre, float r_max,
                            float z_max, float pitch, ckf_bfield* out, float** table);
void tamd_solenoid_map_free(float* table);
/* Exact (un-tabulated) solenoid field at (r, z), tesla; for tests. */
void tamd_solenoid_field_exact(float radius, float length, float b_centre, double r,
                               double z, double* br, double* bz);
/* Bulirsch's generalised complete elliptic integral cel(kc, p, a, b); for tests. */
double tamd_cel(double kc, double p, double a, double b);

typedef enum tamd_pt_distribution {
    TAMD_PT_UNIFORM = 0,     /* U[pt_a, pt_b]                                   */
    TAMD_PT_LOG_UNIFORM = 1, /* log-uniform on [pt_a, pt_b]                      */
    TAMD_PT_SHIFTED_EXP = 2  /* pt_a + Exp(mean pt_b)                            */
} tamd_pt_distribution;

typedef struct tamd_event_spec {
    uint32_t n_particles;
    int32_t pdg;               /* 13 muon, 211 pion (|pdg| used; sign from charge)  */
    float mass;                /* GeV                                               */
    uint32_t random_charge;    /* 1: charge +-1 at random; 0: charge from the pdg sign */
    uint32_t pt_distribution;  /* tamd_pt_distribution                              */
    float pt_a, pt_b;          /* GeV                                               */
    float eta_min, eta_max;
    float phi_min, phi_max;
    float vertex[3];           /* mean vertex, mm                                    */
    float vertex_sigma[3];     /* Gaussian vertex spread, mm                          */
    uint32_t fixed_direction;  /* 1: use dir_phi / dir_theta for every particle      */
    float dir_phi, dir_theta;
    float smear[2];            /* measurement smearing sigma per local coordinate, mm */
    float noise_fraction;      /* noise hits per module ~ Poisson(fraction * signal hits) */
    float dead_fraction;       /* fraction of modules that record no hits              */
    float seed_qop_rel_sigma;  /* seed q/p smeared by (1 + N(0, sigma))                */
    float seed_angle_sigma;    /* seed phi and theta smeared by N(0, sigma), rad       */
    float seed_time_sigma;     /* ns, only enters the seed covariance                  */
    uint32_t max_hits_per_particle;
    uint64_t seed;             /* RNG seed                                             */
} tamd_event_spec;

typedef struct tamd_event {
    /* measurements, sorted by (surface_link, loc0, loc1, var0, var1) */
    uint32_t n_meas;
    float* local_position;   /* [n][2] */
    float* local_variance;   /* [n][2] */
    uint32_t* dimensions;
    float* time;
    float* diameter;
    uint32_t* identifier;
    uint64_t* surface_link;
    uint8_t* subspace;       /* [n][2] */
    uint32_t* cluster_index;
    int32_t* meas_particle;  /* truth particle index, -1 for noise */
    /* truth seeds: one per particle that left at least one measurement */
    uint32_t n_seeds;
    uint64_t* seed_surface_link;
    float* seed_vector;      /* [n][6] */
    float* seed_covariance;  /* [n][36] */
    uint32_t* seed_particle;
    /* particles */
    uint32_t n_particles;
    float* particle_vertex;    /* [n][3] */
    float* particle_momentum;  /* [n][3] */
    float* particle_charge;
    uint32_t* particle_n_meas; /* measurements left in the detector */
} tamd_event;

int tamd_event_generate(const ckf_detector* det, const ckf_bfield* field,
                        const tamd_event_spec* spec, tamd_event* out);
void tamd_event_free(tamd_event* ev);

#ifdef __cplusplus
}
#endif

#endif /* TRACCC_AMD_SIMULATION_H */
