// traccc_amd -- C++ host interface of the MI355X CKF track finder.
//
// Header-only layer over the C ABI in ckf.h that mirrors traccc's operator
// class for this path:
//
//   traccc::host::combinatorial_kalman_filter_algorithm
//     core/include/traccc/finding/combinatorial_kalman_filter_algorithm.hpp:34-72
//   traccc::cuda::combinatorial_kalman_filter_algorithm
//     device/cuda/include/traccc/cuda/finding/combinatorial_kalman_filter_algorithm.hpp:20-211
//
// Same class name, config_type = finding_config, constructor that validates
// the configuration and throws std::invalid_argument where traccc does
// (core/src/finding/combinatorial_kalman_filter_algorithm.cpp:21-32 and
// device/common/src/finding/combinatorial_kalman_filter_algorithm.cpp:33-46),
// and an operator() over (detector, field, measurements, seeds) that returns
// the track candidates.  Inputs and outputs can stay in HBM between calls
// (device_detector / device_events / device_tracks); the host-view overload
// uploads, runs and leaves the result in HBM for download().
//
// Errors other than invalid configuration surface as traccc_amd::error
// (a std::runtime_error carrying the ckf_status).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "traccc_amd/ckf.h"

namespace traccc_amd {

class error : public std::runtime_error {
public:
    error(ckf_status s, const std::string& msg) : std::runtime_error(msg), status_(s) {}
    ckf_status status() const noexcept { return status_; }

private:
    ckf_status status_;
};

inline void check(ckf_status s) {
    if (s == CKF_OK) {
        return;
    }
    const char* m = ckf_last_error();
    const std::string msg = (m != nullptr && *m != '\0') ? m : "traccc_amd error";
    if (s == CKF_ERR_INVALID_ARGUMENT) {
        throw std::invalid_argument(msg);
    }
    throw error(s, msg);
}

// traccc::finding_config (core/include/traccc/finding/finding_config.hpp:70-148)
// with traccc's defaults filled in.
struct finding_config : ckf_finding_config {
    finding_config() : ckf_finding_config() { ckf_finding_config_default(this); }
};

// Geometry + field in HBM (traccc: detector_buffer + device magnetic field).
class device_detector {
public:
    device_detector(const ckf_detector& det, const ckf_bfield& field) {
        ckf_device_detector* h = nullptr;
        check(ckf_device_detector_create(&det, &field, &h));
        h_.reset(h);
    }
    const ckf_device_detector* get() const noexcept { return h_.get(); }

private:
    struct del {
        void operator()(ckf_device_detector* p) const noexcept { ckf_device_detector_destroy(p); }
    };
    std::unique_ptr<ckf_device_detector, del> h_;
};

// One or more events' measurements and seeds in HBM.
class device_events {
public:
    device_events(const device_detector& det, const ckf_measurements& meas,
                  const ckf_bound_params& seeds)
        : device_events(det, &meas, &seeds, 1u) {}
    device_events(const device_detector& det, const ckf_measurements* meas,
                  const ckf_bound_params* seeds, uint32_t n_events) {
        ckf_device_event* h = nullptr;
        check(ckf_device_event_create(det.get(), meas, seeds, n_events, &h));
        h_.reset(h);
    }
    const ckf_device_event* get() const noexcept { return h_.get(); }

private:
    struct del {
        void operator()(ckf_device_event* p) const noexcept { ckf_device_event_destroy(p); }
    };
    std::unique_ptr<ckf_device_event, del> h_;
};

// Host copy of traccc's track container: per-track columns plus the
// measurement indices of each track, in [offsets[t], offsets[t + 1]).
struct track_candidates {
    std::vector<uint16_t> fit_outcome;
    std::vector<uint64_t> params_surface_link;
    std::vector<float> params_vector;      // 6 per track
    std::vector<float> params_covariance;  // 36 per track
    std::vector<float> ndf, chi2, pval;
    std::vector<uint32_t> nholes;
    std::vector<uint32_t> offsets;         // n_tracks + 1
    std::vector<uint16_t> constituent_type;
    std::vector<uint32_t> constituent_index;

    std::size_t size() const noexcept { return chi2.size(); }
};

// Track candidates resident in HBM.
class device_tracks {
public:
    explicit device_tracks(ckf_device_tracks* h) : h_(h) {}

    std::pair<uint32_t, uint32_t> sizes() const {
        uint32_t nt = 0, nc = 0;
        check(ckf_tracks_size(h_.get(), &nt, &nc));
        return {nt, nc};
    }
    std::size_t size() const { return sizes().first; }

    // Per-event track offsets of a batched call (n_events + 1 entries).
    std::vector<uint32_t> event_offsets() const {
        uint32_t ne = 0;
        check(ckf_tracks_event_offsets(h_.get(), &ne, nullptr, nullptr));
        std::vector<uint32_t> off(ne + 1u);
        check(ckf_tracks_event_offsets(h_.get(), &ne, off.data(), nullptr));
        return off;
    }

    track_candidates download() const {
        const auto [nt, nc] = sizes();
        track_candidates t;
        t.fit_outcome.resize(nt);
        t.params_surface_link.resize(nt);
        t.params_vector.resize(6u * nt);
        t.params_covariance.resize(36u * nt);
        t.ndf.resize(nt);
        t.chi2.resize(nt);
        t.pval.resize(nt);
        t.nholes.resize(nt);
        t.offsets.resize(nt + 1u);
        t.constituent_type.resize(nc);
        t.constituent_index.resize(nc);
        ckf_track_candidates out{};
        out.n_tracks = nt;
        out.n_constituents = nc;
        out.fit_outcome = t.fit_outcome.data();
        out.params_surface_link = t.params_surface_link.data();
        out.params_vector = t.params_vector.data();
        out.params_covariance = t.params_covariance.data();
        out.ndf = t.ndf.data();
        out.chi2 = t.chi2.data();
        out.pval = t.pval.data();
        out.nholes = t.nholes.data();
        out.constituent_offsets = t.offsets.data();
        out.constituent_type = t.constituent_type.data();
        out.constituent_index = t.constituent_index.data();
        check(ckf_tracks_download(h_.get(), &out));
        return t;
    }

private:
    struct del {
        void operator()(ckf_device_tracks* p) const noexcept { ckf_tracks_destroy(p); }
    };
    std::unique_ptr<ckf_device_tracks, del> h_;
};

class combinatorial_kalman_filter_algorithm {
public:
    using config_type = finding_config;
    using output_type = device_tracks;

    // hip_stream: a hipStream_t (as void*) every launch is ordered on, or
    // nullptr for the algorithm's own stream.
    explicit combinatorial_kalman_filter_algorithm(const config_type& config,
                                                   void* hip_stream = nullptr) {
        ckf_algorithm* h = nullptr;
        check(ckf_algorithm_create(&config, &h));
        h_.reset(h);
        if (hip_stream != nullptr) {
            check(ckf_algorithm_set_stream(h_.get(), hip_stream));
        }
    }

    // operator() on HBM-resident inputs; the result stays in HBM.
    output_type operator()(const device_detector& det, const device_events& events) const {
        ckf_device_tracks* out = nullptr;
        check(ckf_find_tracks(h_.get(), det.get(), events.get(), &out));
        return output_type(out);
    }

    // operator() on host views of one event (traccc's host signature).
    output_type operator()(const ckf_detector& det, const ckf_bfield& field,
                           const ckf_measurements& measurements,
                           const ckf_bound_params& seeds) const {
        ckf_device_tracks* out = nullptr;
        check(ckf_find_tracks_host(h_.get(), &det, &field, &measurements, &seeds, &out));
        return output_type(out);
    }

    ckf_run_stats last_stats() const {
        ckf_run_stats s{};
        check(ckf_algorithm_last_stats(h_.get(), &s));
        return s;
    }

private:
    struct del {
        void operator()(ckf_algorithm* p) const noexcept { ckf_algorithm_destroy(p); }
    };
    std::unique_ptr<ckf_algorithm, del> h_;
};

}  // namespace traccc_amd
