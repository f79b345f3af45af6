// Exercises the C++ host interface (combinatorial_kalman_filter_algorithm.hpp)
// the way traccc's own tests drive the operator class.
//
//   ckf_cpp_api config          invalid configurations throw std::invalid_argument
//                               (no GPU needed: validation precedes device work)
//   ckf_cpp_api event <n_muons> the same event's sizes and a hash of the
//                               measurement positions (no GPU)
//   ckf_cpp_api run <n_muons>   one toy event (4-layer barrel, 2 T) through both
//                               operator() forms; prints "seeds <s> tracks <n>
//                               constituents <m> hash <h>"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "traccc_amd/combinatorial_kalman_filter_algorithm.hpp"
#include "traccc_amd/simulation.h"

namespace {

int expect_invalid(const char* what, traccc_amd::finding_config cfg) {
    try {
        traccc_amd::combinatorial_kalman_filter_algorithm alg(cfg);
    } catch (const std::invalid_argument& e) {
        std::printf("ok invalid %s: %s\n", what, e.what());
        std::fflush(stdout);
        return 0;
    }
    std::printf("FAIL %s accepted\n", what);
    return 1;
}

int run_config() {
    int bad = 0;
    {
        traccc_amd::finding_config c;
        c.min_track_candidates_per_track = 0;
        bad += expect_invalid("min_track_candidates_per_track=0", c);
    }
    {
        traccc_amd::finding_config c;
        c.min_step_length_for_next_surface = 0.05f;
        c.propagation.overstep_tolerance = -0.1f;
        bad += expect_invalid("min_step_length <= |overstep_tolerance|", c);
    }
    {
        traccc_amd::finding_config c;
        c.run_smoother = CKF_SMOOTHER_KALMAN;
        bad += expect_invalid("kalman smoother", c);
    }
    return bad;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "config") == 0) {
        return run_config() == 0 ? 0 : 1;
    }
    if (argc >= 3 && (std::strcmp(argv[1], "run") == 0 || std::strcmp(argv[1], "event") == 0)) {
        const bool run = std::strcmp(argv[1], "run") == 0;
        tamd_detector sd{};
        if (tamd_detector_build(TAMD_DET_BARREL4, 0.f, 0u, 0.f, &sd) != 0) {
            std::printf("FAIL detector\n");
            return 1;
        }
        const ckf_detector det{sd.surfaces, sd.n_surfaces, sd.layers, sd.n_layers};
        ckf_bfield field{};
        field.kind = CKF_BFIELD_CONSTANT;
        field.b[2] = 2.f;
        // 4-layer barrel muon gun: pT U[1, 10] GeV, |eta| < 1, 50 um smearing
        tamd_event_spec gun{};
        gun.n_particles = static_cast<uint32_t>(std::atoi(argv[2]));
        gun.pdg = 13;
        gun.mass = 0.1056583755f;
        gun.random_charge = 1u;
        gun.pt_distribution = TAMD_PT_UNIFORM;
        gun.pt_a = 1.f;
        gun.pt_b = 10.f;
        gun.eta_min = -1.f;
        gun.eta_max = 1.f;
        gun.phi_min = -3.14159265f;
        gun.phi_max = 3.14159265f;
        gun.vertex_sigma[0] = 0.1f;
        gun.vertex_sigma[1] = 0.1f;
        gun.vertex_sigma[2] = 50.f;
        gun.dir_theta = 1.57079637f;
        gun.smear[0] = gun.smear[1] = 0.05f;
        gun.seed_qop_rel_sigma = 0.1f;
        gun.seed_angle_sigma = 0.001f;
        gun.seed_time_sigma = 1.f;
        gun.max_hits_per_particle = 64u;
        gun.seed = 42u;
        tamd_event ev{};
        if (tamd_event_generate(&det, &field, &gun, &ev) != 0) {
            std::printf("FAIL event\n");
            return 1;
        }
        const ckf_measurements meas{ev.n_meas, ev.local_position, ev.local_variance,
                                    ev.dimensions, ev.time, ev.diameter, ev.identifier,
                                    ev.surface_link, ev.subspace, ev.cluster_index};
        const ckf_bound_params seeds{ev.n_seeds, ev.seed_surface_link, ev.seed_vector,
                                     ev.seed_covariance};
        if (!run) {
            uint64_t h = 1469598103934665603ull;
            for (uint32_t i = 0; i < 2u * ev.n_meas; ++i) {
                uint32_t bits;
                std::memcpy(&bits, &ev.local_position[i], 4);
                h = (h ^ bits) * 1099511628211ull;
            }
            std::printf("seeds %u measurements %u hash %016llx\n", ev.n_seeds, ev.n_meas,
                        static_cast<unsigned long long>(h));
            tamd_event_free(&ev);
            tamd_detector_free(&sd);
            return 0;
        }

        // muon hypothesis, pT floor 0.2 GeV, no smoother
        traccc_amd::finding_config cfg;
        cfg.ptc_pdg = 13;
        cfg.ptc_mass = 0.1056583755f;
        cfg.ptc_charge = -1.f;
        cfg.min_pT = 0.2f;
        cfg.run_smoother = CKF_SMOOTHER_NONE;
        traccc_amd::combinatorial_kalman_filter_algorithm alg(cfg);
        // host-view operator and the HBM-resident form must agree
        const auto t = alg(det, field, meas, seeds).download();
        const traccc_amd::device_detector ddet(det, field);
        const traccc_amd::device_events devs(ddet, meas, seeds);
        const auto t2 = alg(ddet, devs).download();
        if (t2.constituent_index != t.constituent_index || t2.offsets != t.offsets) {
            std::printf("FAIL host-view and device-resident results differ\n");
            return 1;
        }
        uint64_t h = 1469598103934665603ull;
        for (uint32_t v : t.constituent_index) { h = (h ^ v) * 1099511628211ull; }
        for (uint32_t v : t.offsets) { h = (h ^ v) * 1099511628211ull; }
        std::printf("seeds %u tracks %zu constituents %zu hash %016llx\n", ev.n_seeds, t.size(),
                    t.constituent_index.size(), static_cast<unsigned long long>(h));
        tamd_event_free(&ev);
        tamd_detector_free(&sd);
        return 0;
    }
    std::fprintf(stderr, "usage: ckf_cpp_api config | event <n_muons> | run <n_muons>\n");
    return 2;
}
