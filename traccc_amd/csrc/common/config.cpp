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
