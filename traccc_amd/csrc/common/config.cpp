// finding_config defaults and validation.
//
// Defaults follow traccc::finding_config (core/include/traccc/finding/finding_config.hpp:72-146).
// Propagation defaults are this project's navigator settings, named after the
// detray::propagation::config members they stand in for.
#include <cmath>
#include <cstring>
#include <limits>

#include "traccc_amd/ckf.h"
#include "traccc_amd/fitting.h"
#include "traccc_amd/seeding.h"

extern "C" void ckf_finding_config_default(ckf_finding_config* c) {
    if (c == nullptr) {
        return;
    }
    std::memset(c, 0, sizeof(*c));
    c->max_num_branches_per_seed = 10u;
    c->max_num_branches_per_surface = 1u;
    c->min_track_candidates_per_track = 3u;
    c->max_track_candidates_per_track = 100u;
    c->max_num_skipping_per_cand = 3u;
    c->max_num_consecutive_skipped = 1u;
    c->max_num_tracks_per_measurement = 0u;
    c->min_measurement_voting_fraction = 0.5f;
    c->run_pkf = 0u;
    c->run_smoother = CKF_SMOOTHER_MBF;
    c->min_step_length_for_next_surface = 1.2f;
    c->max_step_counts_for_next_surface = 100u;
    c->chi2_max = 30.f;
    c->propagation.step_constraint = std::numeric_limits<float>::max();
    c->propagation.rk_error_tol = 1e-4f;
    c->propagation.min_stepsize = 1e-4f;
    c->propagation.path_limit = 5000.f;
    c->propagation.overstep_tolerance = -0.1f;
    c->propagation.min_mask_tolerance = 0.5f;
    c->propagation.max_mask_tolerance = 3.f;
    c->propagation.mask_tolerance_scalor = 0.05f;
    c->propagation.on_surface_tolerance = 1e-3f;
    c->propagation.search_window[0] = 1u;
    c->propagation.search_window[1] = 1u;
    c->min_p = 0.1f;
    c->min_pT = 0.6f;
    c->ptc_pdg = 211;
    c->ptc_mass = 0.13957039f;
    c->ptc_charge = 1.f;
    c->duplicate_removal_minimum_length = 5u;
    c->initial_links_per_seed = 100u;
}

// fitting_config defaults: traccc::fitting_config
// (core/include/traccc/fitting/fitting_config.hpp:20-45); the propagation,
// momentum cuts and particle hypothesis default to finding_config's.
extern "C" void ckf_fitting_config_default(ckf_fitting_config* cfg) {
    if (cfg == nullptr) {
        return;
    }
    ckf_finding_config f{};
    ckf_finding_config_default(&f);
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->n_iterations = 1u;
    cfg->propagation = f.propagation;
    cfg->max_step_counts = f.max_step_counts_for_next_surface;
    cfg->min_p = f.min_p;
    cfg->min_pT = f.min_pT;
    cfg->ptc_pdg = f.ptc_pdg;
    cfg->ptc_mass = f.ptc_mass;
    cfg->ptc_charge = f.ptc_charge;
    cfg->covariance_inflation_factor = 1e3f;
    cfg->surface_sequence_size_factor = 5u;
    cfg->min_surface_sequence_capacity = 100u;
    cfg->do_precise_hole_count = 0u;
}

// seedfinder_config defaults (core/include/traccc/seeding/detail/seeding_config.hpp:21-110),
// in traccc units: mm, GeV, field in GeV / (e mm) (1 T = 0.000299792458).
extern "C" void ckf_seedfinder_config_default(ckf_seedfinder_config* c) {
    if (c == nullptr) {
        return;
    }
    std::memset(c, 0, sizeof(*c));
    c->zMin = -2000.f;
    c->zMax = 2000.f;
    c->rMax = 200.f;
    c->rMin = 33.f;
    c->collisionRegionMin = -250.f;
    c->collisionRegionMax = 250.f;
    c->phiMin = static_cast<float>(-M_PI);
    c->phiMax = static_cast<float>(M_PI);
    c->minPt = 500.f * 0.001f;
    c->cotThetaMax = 27.2845f;
    c->deltaRMin = 20.f;
    c->deltaRMax = 80.f;
    c->deltaZMax = 450.f;
    c->impactMax = 10.f;
    c->sigmaScattering = 3.0f;
    c->maxPtScattering = 10.f;
    c->maxSeedsPerSpM = 5u;
    c->bFieldInZ = 1.99724f * 0.000299792458f;
    c->radLengthPerSeed = 0.05f;
    c->sigmaError = 5.f;
    c->phiBinDeflectionCoverage = 1;
    c->neighbor_scope[0] = 1u;
    c->neighbor_scope[1] = 1u;
    ckf_seedfinder_config_setup(c);
}

// seedfinder_config::setup (seeding_config.hpp:119-134); MeV = 0.001 GeV
extern "C" void ckf_seedfinder_config_setup(ckf_seedfinder_config* c) {
    if (c == nullptr) {
        return;
    }
    c->highland = 13.6f * 0.001f * std::sqrt(c->radLengthPerSeed) *
                  (1.f + 0.038f * std::log(c->radLengthPerSeed));
    const float maxScatteringAngle = c->highland / c->minPt;
    c->maxScatteringAngle2 = maxScatteringAngle * maxScatteringAngle;
    c->pTPerHelixRadius = c->bFieldInZ;
    c->minHelixDiameter2 = std::pow(c->minPt * 2.f / c->pTPerHelixRadius, 2.f);
    c->minHelixRadius = std::sqrt(c->minHelixDiameter2) / 2.f;
    c->pT2perRadius = std::pow(c->highland / c->pTPerHelixRadius, 2.f);
}

// spacepoint_grid_config(const seedfinder_config&) (seeding_config.hpp:147-160)
extern "C" void ckf_spacepoint_grid_config_from(const ckf_seedfinder_config* f,
                                                ckf_spacepoint_grid_config* g) {
    if (f == nullptr || g == nullptr) {
        return;
    }
    g->bFieldInZ = f->bFieldInZ;
    g->minPt = f->minPt;
    g->rMax = f->rMax;
    g->zMax = f->zMax;
    g->zMin = f->zMin;
    g->deltaRMax = f->deltaRMax;
    g->cotThetaMax = f->cotThetaMax;
    g->impactMax = f->impactMax;
    g->phiMin = f->phiMin;
    g->phiMax = f->phiMax;
    g->phiBinDeflectionCoverage = f->phiBinDeflectionCoverage;
}

// seedfilter_config defaults (seeding_config.hpp:202-238)
extern "C" void ckf_seedfilter_config_default(ckf_seedfilter_config* c) {
    if (c == nullptr) {
        return;
    }
    c->deltaInvHelixDiameter = 0.00003f;
    c->impactWeightFactor = 1.f;
    c->compatSeedWeight = 200.f;
    c->deltaRMin = 5.f;
    c->compatSeedLimit = 2u;
    c->good_spB_min_radius = 150.f;
    c->good_spB_weight_increase = 400.f;
    c->good_spT_max_radius = 150.f;
    c->good_spT_weight_increase = 200.f;
    c->good_spB_min_weight = 380.f;
    c->seed_min_weight = 200.f;
    c->spB_min_radius = 43.f;
}

// track_params_estimation_config defaults
// (core/include/traccc/seeding/detail/track_params_estimation_config.hpp:18-31);
// 1 degree = pi / 180, 1 ns = 1 (traccc's time unit)
extern "C" void ckf_track_params_estimation_config_default(ckf_track_params_estimation_config* c) {
    if (c == nullptr) {
        return;
    }
    const float deg = static_cast<float>(M_PI / 180.0);
    const float sig[6] = {1.f, 1.f, deg, deg, 0.f, 1.f};
    const float infl[6] = {1.f, 1.f, 1.f, 1.f, 1.f, 100.f};
    for (int i = 0; i < 6; ++i) {
        c->initial_sigma[i] = sig[i];
        c->initial_inflation[i] = infl[i];
    }
    c->initial_sigma_qopt = 0.1f;
    c->initial_sigma_pt_rel = 0.1f;
}
