// Host code of the project under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by tests/test_sanitizers.py with g++,
// no device code): the CPU oracle (CKF in FP64 and FP32, the Kalman fit,
// the greedy ambiguity resolution),
// the synthetic-input toolkit (detectors, solenoid map, particle gun) and the
// CSV measurement reader / writer, including malformed files -- the reader
// parses text from outside the program.  Exits 0 when every call behaved;
// a sanitizer finding aborts the run with its report on stderr.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"
#include "traccc_amd/io.h"
#include "traccc_amd/simulation.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        ++failures;
    }
}

ckf_measurements meas_view(const tamd_event& ev) {
    ckf_measurements m{};
    m.size = ev.n_meas;
    m.local_position = ev.local_position;
    m.local_variance = ev.local_variance;
    m.dimensions = ev.dimensions;
    m.time = ev.time;
    m.diameter = ev.diameter;
    m.identifier = ev.identifier;
    m.surface_link = ev.surface_link;
    m.subspace = ev.subspace;
    m.cluster_index = ev.cluster_index;
    return m;
}

ckf_bound_params seed_view(const tamd_event& ev) {
    ckf_bound_params s{};
    s.size = ev.n_seeds;
    s.surface_link = ev.seed_surface_link;
    s.vector = ev.seed_vector;
    s.covariance = ev.seed_covariance;
    return s;
}

void run_oracle(uint32_t preset, float x0, bool map, uint32_t n_particles, uint32_t branches) {
    tamd_detector det{};
    expect(tamd_detector_build(preset, x0, 9, 20.f, &det) == 0, "detector");
    ckf_detector d{det.surfaces, det.n_surfaces, det.layers, det.n_layers};
    ckf_bfield field{};
    float* table = nullptr;
    if (map) {
        expect(tamd_solenoid_map_build(1200.f, 5600.f, 2.f, 1500.f, 3000.f, 10.f, &field, &table) == 0,
               "solenoid map");
    } else {
        field.kind = CKF_BFIELD_CONSTANT;
        field.b[2] = 2.f;
    }
    tamd_event_spec spec{};
    spec.n_particles = n_particles;
    spec.pdg = 211;
    spec.mass = 0.13957039f;
    spec.random_charge = 1;
    spec.pt_distribution = TAMD_PT_SHIFTED_EXP;
    spec.pt_a = 0.3f;
    spec.pt_b = 1.f;
    spec.eta_min = -2.5f;
    spec.eta_max = 2.5f;
    spec.phi_min = -3.14159265f;
    spec.phi_max = 3.14159265f;
    spec.vertex_sigma[0] = spec.vertex_sigma[1] = 0.1f;
    spec.vertex_sigma[2] = 50.f;
    spec.smear[0] = spec.smear[1] = 0.05f;
    spec.noise_fraction = 0.1f;
    spec.dead_fraction = 0.03f;
    spec.seed_qop_rel_sigma = 0.1f;
    spec.seed_angle_sigma = 0.001f;
    spec.seed_time_sigma = 1.f;
    spec.max_hits_per_particle = 64;
    spec.seed = 42;
    tamd_event ev{};
    expect(tamd_event_generate(&d, &field, &spec, &ev) == 0, "event");
    ckf_finding_config cfg;
    ckf_finding_config_default(&cfg);
    cfg.max_num_branches_per_surface = branches;
    cfg.ptc_pdg = 211;
    cfg.ptc_mass = 0.13957039f;
    cfg.ptc_charge = 1.f;
    cfg.min_pT = 0.2f;
    cfg.run_smoother = CKF_SMOOTHER_NONE;
    const ckf_measurements m = meas_view(ev);
    const ckf_bound_params s = seed_view(ev);
    for (int precision : {64, 32}) {
        oracle_tracks t{};
        expect(oracle_ckf_run(&cfg, &d, &field, &m, &s, precision, &t) == 0, "oracle ckf");
        if (precision == 64 && t.n_tracks > 0) {
            // the Kalman fit of the found candidates
            ckf_fitting_config fc;
            ckf_fitting_config_default(&fc);
            fc.ptc_pdg = 211;
            fc.ptc_mass = 0.13957039f;
            fc.ptc_charge = 1.f;
            fc.min_pT = 0.2f;
            oracle_fitted f{};
            double secs = 0.0;
            expect(oracle_fit_run(&fc, &d, &field, &m, t.n_tracks, t.params_surface_link,
                                  t.params_vector, t.params_covariance, t.constituent_offsets,
                                  t.constituent_index, 64, 2, &f, &secs) == 0,
                   "oracle fit");
            oracle_fitted_free(&f);
        }
        oracle_tracks_free(&t);
    }
    // measurement-vote pruning path of the oracle
    cfg.max_num_tracks_per_measurement = 2;
    oracle_tracks t{};
    expect(oracle_ckf_run(&cfg, &d, &field, &m, &s, 64, &t) == 0, "oracle votes");
    oracle_tracks_free(&t);
    tamd_event_free(&ev);
    if (table) tamd_solenoid_map_free(table);
    tamd_detector_free(&det);
}

// Greedy ambiguity resolution on random track sets: duplicated identifiers,
// empty tracks, every configuration path.
void run_ambiguity() {
    uint64_t rng = 12345;
    auto next = [&rng](uint32_t mod) {
        rng = rng * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<uint32_t>((rng >> 33) % mod);
    };
    for (int rep = 0; rep < 6; ++rep) {
        const uint32_t n = 50 + next(400);
        std::vector<uint32_t> off(n + 1, 0), ids;
        std::vector<float> pv(n);
        for (uint32_t t = 0; t < n; ++t) {
            const uint32_t len = next(12);
            for (uint32_t j = 0; j < len; ++j) ids.push_back(next(rep % 2 ? 60 : 600));
            off[t + 1] = static_cast<uint32_t>(ids.size());
            pv[t] = static_cast<float>(next(1000)) / 1000.f;
        }
        std::vector<uint8_t> acc(n);
        const uint32_t cfgs[4][3] = {{3, 0xffffffffu, 1}, {0, 0xffffffffu, 1}, {3, 7, 1}, {2, 0xffffffffu, 2}};
        for (const auto& c : cfgs) {
            uint32_t removed = 0;
            expect(oracle_greedy_resolve(n, off.data(), ids.data(), pv.data(), c[0], c[1], c[2],
                                         acc.data(), &removed) == 0,
                   "greedy resolve");
        }
    }
    uint8_t a = 0;
    expect(oracle_greedy_resolve(0, nullptr, nullptr, nullptr, 3, 10, 1, &a, nullptr) == 0, "empty");
    expect(oracle_greedy_resolve(0, nullptr, nullptr, nullptr, 3, 10, 0, &a, nullptr) != 0,
           "zero threshold rejected");
}

void write_file(const std::string& path, const std::string& text) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        expect(false, "open temp file");
        return;
    }
    std::fwrite(text.data(), 1, text.size(), f);
    std::fclose(f);
}

void run_csv(const char* mock, const std::string& dir) {
    ckf_measurement_buffer* b = nullptr;
    expect(ckf_read_measurements_csv(mock, 1, &b) == CKF_OK, "read mock csv");
    ckf_measurements v{};
    const uint32_t* idx = nullptr;
    expect(ckf_measurement_buffer_view(b, &v, &idx) == CKF_OK, "view");
    const std::string out = dir + "/roundtrip.csv";
    expect(ckf_write_measurements_csv(out.c_str(), &v) == CKF_OK, "write csv");
    ckf_measurement_buffer* b2 = nullptr;
    expect(ckf_read_measurements_csv(out.c_str(), 0, &b2) == CKF_OK, "read back");
    ckf_measurements v2{};
    expect(ckf_measurement_buffer_view(b2, &v2, nullptr) == CKF_OK, "view back");
    expect(v2.size == v.size, "round-trip size");
    ckf_measurement_buffer_destroy(b2);
    ckf_measurement_buffer_destroy(b);

    const std::string head =
        "measurement_id,geometry_id,local_key,local0,local1,phi,theta,time,var_local0,"
        "var_local1,var_phi,var_theta,var_time\n";
    const std::vector<std::string> bad = {
        "",                                               // empty file
        "\n\n",                                           // no header
        "geometry_id,local0\n1,2\n",                      // missing columns
        head + "0,1,6,0.1\n",                             // truncated row
        head + "0,1,6,abc,0.2,0,0,0,0.0025,0.0025,0,0,0\n",   // unparsable number
        head + "0,1,0,0.1,0.2,0,0,0,0.0025,0.0025,0,0,0\n",   // local_key measures nothing
        head + "0,1,6,1e999999,0.2,0,0,0,0.0025,0.0025,0,0,0\n",
        head + "0,18446744073709551616,6,0.1,0.2,0,0,0,0.0025,0.0025,0,0,0\n",
        head + "0,1,6,0.1,0.2,0,0,0,0.0025,0.0025,0,0,0,extra,columns,,,\n",
        head + std::string(100000, ','),                  // one huge row of empty fields
        std::string("\xff\xfe\x00\x01", 4) + head,        // binary junk before the header
        head + "0,1,6,0.1,0.2\r\n",                       // CRLF, short
    };
    int k = 0;
    for (const std::string& text : bad) {
        const std::string path = dir + "/bad" + std::to_string(k++) + ".csv";
        write_file(path, text);
        ckf_measurement_buffer* bb = nullptr;
        const ckf_status st = ckf_read_measurements_csv(path.c_str(), 1, &bb);
        // malformed input is an error (or, for inputs a lenient parse
        // accepts, a valid buffer): never a crash
        if (st == CKF_OK) {
            ckf_measurements vv{};
            expect(ckf_measurement_buffer_view(bb, &vv, nullptr) == CKF_OK, "view of accepted");
            ckf_measurement_buffer_destroy(bb);
        } else {
            expect(bb == nullptr, "no buffer on error");
            expect(std::strlen(ckf_last_error()) > 0, "error message");
        }
    }
    ckf_measurement_buffer* none = nullptr;
    expect(ckf_read_measurements_csv((dir + "/does-not-exist.csv").c_str(), 0, &none) != CKF_OK,
           "missing file is an error");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <mock measurements csv> <scratch dir>\n", argv[0]);
        return 2;
    }
    run_oracle(TAMD_DET_BARREL4, 0.f, false, 200, 1);     // C1-like: no material
    run_oracle(TAMD_DET_BARREL4_DISK4, 0.01f, false, 300, 3);
    run_oracle(TAMD_DET_BARREL5_DISK6, 0.02f, true, 300, 3);
    run_oracle(TAMD_DET_TELESCOPE, 0.005f, false, 20, 2);
    run_ambiguity();
    run_csv(argv[1], argv[2]);
    std::printf("sanitized run: %d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
