// Handles of the C ABI shared by the CKF and the fitter runtimes: device
// buffers, the per-call result block cache, and the device-resident
// detector, events and track candidates.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "../common/errors.hpp"
#include "traccc_amd/ckf.h"
#include "../core/geometry.hpp"

namespace tamd {

// Device allocation that grows by doubling and is reused across calls.
struct dbuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t need) {
        if (need <= bytes) {
            return;
        }
        if (p != nullptr) {
            tamd::hip_check(hipFree(p), "hipFree");
            p = nullptr;
        }
        size_t nb = std::max<size_t>(need, bytes * 2);
        nb = (nb + 255) & ~size_t(255);
        tamd::hip_check(hipMalloc(&p, nb), "hipMalloc");
        bytes = nb;
    }
    void release() {
        if (p != nullptr) {
            (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Process-wide cache of device blocks for per-call results: freed blocks are
// kept by power-of-two size class and handed out again, so a steady stream of
// operator calls does not call hipMalloc / hipFree.  At most kMaxHeld bytes
// are kept (a returned block beyond that is freed); release() frees them all
// (ckf_release_cached_memory).
struct block_cache {
    static constexpr size_t kMaxHeld = size_t(4) << 30;
    std::mutex mu;
    std::map<size_t, std::vector<void*>> free_blocks;
    size_t held = 0;
    static size_t size_class(size_t bytes) {
        size_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    void* get(size_t bytes) {
        const size_t c = size_class(std::max<size_t>(bytes, 1));
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = free_blocks.find(c);
            if (it != free_blocks.end() && !it->second.empty()) {
                void* p = it->second.back();
                it->second.pop_back();
                held -= c;
                return p;
            }
        }
        void* p = nullptr;
        tamd::hip_check(hipMalloc(&p, c), "hipMalloc");
        return p;
    }
    void put(void* p, size_t bytes) {
        if (p == nullptr) return;
        const size_t c = size_class(std::max<size_t>(bytes, 1));
        std::lock_guard<std::mutex> lk(mu);
        if (held + c > kMaxHeld) {
            (void)hipFree(p);
            return;
        }
        free_blocks[c].push_back(p);
        held += c;
    }
    void release() {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& kv : free_blocks) {
            for (void* p : kv.second) (void)hipFree(p);
        }
        free_blocks.clear();
        held = 0;
    }
};

block_cache& result_cache();

template <typename T>
T* dev_copy(const T* src, size_t n) {
    T* d = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    tamd::hip_check(hipMalloc(&d, bytes), "hipMalloc");
    if (n > 0 && src != nullptr) {
        tamd::hip_check(hipMemcpy(d, src, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    return d;
}

}  // namespace tamd

struct ckf_device_detector {
    std::vector<ckf_surface> h_sfs;
    std::vector<ckf_layer> h_layers;
    std::vector<uint64_t> ids;
    ckf_bfield field{};
    ckf_surface* d_sfs = nullptr;
    ckf_layer* d_layers = nullptr;
    // the navigator's layer groups (tamd::build_layer_groups)
    std::vector<tamd::layer_group> h_groups;
    tamd::layer_group* d_groups = nullptr;
    float* d_map = nullptr;
    ~ckf_device_detector() {
        if (d_sfs) (void)hipFree(d_sfs);
        if (d_groups) (void)hipFree(d_groups);
        if (d_layers) (void)hipFree(d_layers);
        if (d_map) (void)hipFree(d_map);
    }
};

struct ckf_device_event {
    uint32_t n_events = 0, n_meas = 0, n_seeds = 0, n_surfaces = 0;
    std::vector<uint32_t> meas_off, seed_off;  // [n_events + 1]
    float2* d_pos = nullptr;
    float2* d_var = nullptr;
    uint32_t* d_info = nullptr;
    uint32_t* d_ranges = nullptr;
    uint32_t* d_seed_event = nullptr;
    uint32_t* d_seed_surface = nullptr;
    float* d_seed_vec = nullptr;
    float* d_seed_cov = nullptr;
    uint32_t* d_meas_off = nullptr;
    uint32_t* d_meas_sf = nullptr;  // surface index of every measurement
    float* d_win_var = nullptr;     // per event: largest loc0 variance, or -1 (no window search)
    ~ckf_device_event() {
        void* ps[] = {d_pos, d_var, d_info, d_ranges, d_seed_event, d_seed_surface,
                      d_seed_vec, d_seed_cov, d_meas_off, d_meas_sf, d_win_var};
        for (void* p : ps) {
            if (p) (void)hipFree(p);
        }
    }
};

// Fitted tracks in HBM (ckf_fit_tracks): per-track columns, the
// candidates' constituents (event-relative measurement indices) and the
// per-state filtered / smoothed records.
struct ckf_device_fitted {
    uint32_t n_tracks = 0, n_states = 0;
    std::vector<uint64_t> ids;            // surface index -> identifier
    std::vector<uint32_t> ev_track_off;   // [n_events + 1] tracks of each event
    tamd::dbuf outcome, surface, vec, cov, ndf, chi2, pval, nholes, offsets, cons, flt, smt, flags,
        state_sf;
    ~ckf_device_fitted() {
        tamd::dbuf* bs[] = {&outcome, &surface, &vec, &cov, &ndf, &chi2, &pval, &nholes,
                            &offsets, &cons, &flt, &smt, &flags, &state_sf};
        for (tamd::dbuf* b : bs) b->release();
    }
};

struct ckf_device_tracks {
    uint32_t n_events = 0, n_tracks = 0, n_const = 0;
    std::vector<uint32_t> ev_track_off, ev_const_off;
    std::vector<uint64_t> ids;
    uint16_t* outcome = nullptr;
    uint32_t* surface = nullptr;
    float* vec = nullptr;
    float* cov = nullptr;
    float* ndf = nullptr;
    float* chi2 = nullptr;
    float* pval = nullptr;
    uint32_t* nholes = nullptr;
    uint32_t* offsets = nullptr;
    uint32_t* cons = nullptr;
    size_t track_bytes[10] = {};
    ~ckf_device_tracks() {
        void* ps[] = {outcome, surface, vec, cov, ndf, chi2, pval, nholes, offsets, cons};
        for (int i = 0; i < 10; ++i) tamd::result_cache().put(ps[i], track_bytes[i]);
    }
};

