// CPU oracle of the greedy ambiguity resolution -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/ and bench.py's cpu_baseline leg load it (liboracle_ckf.so), as
// the checker and the timed CPU baseline; the product never links it.
//
// This project's restatement of traccc's host algorithm
// (core/src/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.cpp:19-254):
//   * tracks with fewer than min_meas_per_track constituents are rejected
//     up front (:48-73) and take no part in the sharing counts;
//   * a track's measurements are identified by the measurement identifier
//     (:65-69); tracks_per_measurement holds each accepted track once per
//     distinct identifier (:95-109);
//   * n_shared counts the track's constituents (duplicates included) whose
//     identifier is held by more than one accepted track (:111-124),
//     rel_shared = n_shared / n_meas in single precision (:126-131);
//   * the tracks are ordered by rel_shared ascending, then p-value descending
//     (:136-143); each iteration stops when the largest n_shared over the
//     accepted tracks is below max_shared_meas (:152-162), else removes the
//     last (worst) track (:164-174) and, for each distinct identifier of it,
//     drops it from that identifier's tracks; when one track is left there,
//     that track loses the identifier's occurrences from n_shared and moves
//     in the order (:176-227); at most max_iterations iterations (:146).
//
// Where this differs from the reference, on purpose:
//   * exact ties in (rel_shared, p-value) are broken by track index, the
//     higher index being the worse; the reference's order there is whatever
//     std::sort leaves (unspecified for equal keys);
//   * a NaN p-value orders as the lowest p-value; a track with no
//     constituents (possible only with min_meas_per_track = 0) has
//     rel_shared 0 instead of 0/0.
// The order is kept in a balanced tree (std::set) instead of a sorted vector
// with repositioning; with a strict total order both give the same worst
// track at every iteration.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <set>
#include <vector>

#include "oracle.h"

namespace {

struct track_order {
    const std::vector<float>* rel;
    const std::vector<float>* pkey;
    // a < b: a is the better track (sorts earlier)
    bool operator()(uint32_t a, uint32_t b) const {
        const float ra = (*rel)[a], rb = (*rel)[b];
        if (ra != rb) return ra < rb;
        const float pa = (*pkey)[a], pb = (*pkey)[b];
        if (pa != pb) return pa > pb;
        return a < b;
    }
};

}  // namespace

extern "C" int oracle_greedy_resolve(uint32_t n_tracks, const uint32_t* offsets,
                                     const uint32_t* meas_ids, const float* pval,
                                     uint32_t min_meas_per_track, uint32_t max_iterations,
                                     uint32_t max_shared_meas, uint8_t* accepted,
                                     uint32_t* n_removed) {
    if (accepted == nullptr || (n_tracks > 0 && (offsets == nullptr || pval == nullptr)) ||
        max_shared_meas == 0u) {
        return 1;
    }
    uint32_t removed = 0;
    std::memset(accepted, 0, n_tracks);
    if (n_removed) *n_removed = 0;
    if (n_tracks == 0) return 0;

    std::vector<float> rel(n_tracks, 0.f), pkey(n_tracks);
    std::vector<uint32_t> n_meas(n_tracks), n_shared(n_tracks, 0u);
    for (uint32_t t = 0; t < n_tracks; ++t) {
        n_meas[t] = offsets[t + 1] - offsets[t];
        pkey[t] = std::isnan(pval[t]) ? -std::numeric_limits<float>::infinity() : pval[t];
        accepted[t] = n_meas[t] >= min_meas_per_track ? 1u : 0u;
    }
    // distinct (identifier, track) pairs of the accepted tracks, grouped by
    // identifier: tracks_per_measurement
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint32_t t = 0; t < n_tracks; ++t) {
        if (!accepted[t]) continue;
        for (uint32_t j = offsets[t]; j < offsets[t + 1]; ++j) pairs.emplace_back(meas_ids[j], t);
    }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    std::vector<uint32_t> ids;      // distinct identifiers, ascending
    std::vector<uint32_t> grp_off;  // [ids.size() + 1] into pairs
    for (size_t k = 0; k < pairs.size(); ++k) {
        if (k == 0 || pairs[k].first != pairs[k - 1].first) {
            ids.push_back(pairs[k].first);
            grp_off.push_back(static_cast<uint32_t>(k));
        }
    }
    grp_off.push_back(static_cast<uint32_t>(pairs.size()));
    auto group_of = [&](uint32_t id) {
        return static_cast<size_t>(std::lower_bound(ids.begin(), ids.end(), id) - ids.begin());
    };
    std::vector<uint32_t> alive_on(ids.size());  // accepted tracks still holding the identifier
    for (size_t g = 0; g < ids.size(); ++g) alive_on[g] = grp_off[g + 1] - grp_off[g];

    auto rel_of = [&](uint32_t t) {
        return n_meas[t] == 0u ? 0.f
                               : static_cast<float>(n_shared[t]) / static_cast<float>(n_meas[t]);
    };
    for (uint32_t t = 0; t < n_tracks; ++t) {
        if (!accepted[t]) continue;
        for (uint32_t j = offsets[t]; j < offsets[t + 1]; ++j) {
            if (alive_on[group_of(meas_ids[j])] > 1u) ++n_shared[t];
        }
        rel[t] = rel_of(t);
    }
    track_order cmp{&rel, &pkey};
    std::set<uint32_t, track_order> order(cmp);
    // histogram of n_shared over the accepted tracks: its top is max_shared
    uint32_t top = 0;
    std::vector<uint32_t> hist;
    for (uint32_t t = 0; t < n_tracks; ++t) {
        if (!accepted[t]) continue;
        order.insert(t);
        if (n_shared[t] >= hist.size()) hist.resize(n_shared[t] + 1u, 0u);
        ++hist[n_shared[t]];
        top = std::max(top, n_shared[t]);
    }
    std::vector<uint32_t> seen;
    for (uint32_t iter = 0; iter < max_iterations; ++iter) {
        if (order.empty()) break;
        while (top > 0u && hist[top] == 0u) --top;
        if (top < max_shared_meas) break;
        const uint32_t worst = *order.rbegin();
        order.erase(std::prev(order.end()));
        accepted[worst] = 0u;
        --hist[n_shared[worst]];
        ++removed;
        seen.clear();
        for (uint32_t j = offsets[worst]; j < offsets[worst + 1]; ++j) {
            const uint32_t id = meas_ids[j];
            if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
            seen.push_back(id);
            const size_t g = group_of(id);
            if (--alive_on[g] != 1u) continue;
            // one accepted track is left on the identifier: it no longer shares it
            uint32_t tid = 0;
            for (uint32_t k = grp_off[g]; k < grp_off[g + 1]; ++k) {
                if (accepted[pairs[k].second]) {
                    tid = pairs[k].second;
                    break;
                }
            }
            uint32_t cnt = 0;
            for (uint32_t q = offsets[tid]; q < offsets[tid + 1]; ++q) cnt += meas_ids[q] == id ? 1u : 0u;
            order.erase(tid);
            --hist[n_shared[tid]];
            n_shared[tid] -= cnt;
            ++hist[n_shared[tid]];
            rel[tid] = rel_of(tid);
            order.insert(tid);
        }
    }
    if (n_removed) *n_removed = removed;
    return 0;
}
