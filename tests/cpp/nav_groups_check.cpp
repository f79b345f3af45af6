// Host harness: the navigator's layer search over layer groups against the
// plain scan of every layer, on caller-supplied straight-line states.  Both
// run the same float code the device runs (core/propagator.hpp).
#include <cstdint>

#include "core/propagator.hpp"

extern "C" int nav_groups_check(const ckf_layer* layers, uint32_t n_layers, uint32_t n,
                                const float* pos, const float* dir, const uint64_t* visited,
                                float margin, int32_t* best_scan, float* dist_scan,
                                int32_t* best_grouped, float* dist_grouped) {
    tamd::layer_group groups[tamd::kMaxGroups];
    const uint32_t n_groups = tamd::build_layer_groups(layers, n_layers, groups);
    // the boxes as the transport kernels stage them: layers, then the groups at 64
    tamd::nav_box boxes[64 + tamd::kMaxGroups];
    for (uint32_t k = 0; k < n_layers + tamd::kMaxGroups; ++k) {
        boxes[k < n_layers ? k : 64u + (k - n_layers)] =
            tamd::nav_box_of(layers, n_layers, groups, n_groups, margin, k);
    }
    tamd::navigator<float, true, 1> nav{};
    nav.layers = layers;
    nav.n_layers = n_layers;
    nav.cfg.mask_tol_max = margin;
    for (uint32_t i = 0; i < n; ++i) {
        nav.st.pos = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        nav.st.dir = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
        nav.st.lo = {0.f, 0.f, 0.f};
        int best = -1;
        float bs = 0.f;
        nav.groups = nullptr;
        nav.n_groups = 0u;
        nav.layer_boxes = nullptr;
        nav.group_boxes = nullptr;
        nav.pick_layer(visited[i], best, bs);
        best_scan[i] = best;
        dist_scan[i] = bs;
        nav.groups = groups;
        nav.n_groups = n_groups;
        nav.layer_boxes = boxes;
        nav.group_boxes = boxes + 64;
        nav.pick_layer(visited[i], best, bs);
        best_grouped[i] = best;
        dist_grouped[i] = bs;
    }
    return static_cast<int>(n_groups);
}

// The groups themselves: members (-1 past a group's size) and boxes.
extern "C" uint32_t nav_groups_build(const ckf_layer* layers, uint32_t n_layers, int32_t* member,
                                     float* box, uint32_t* bounded) {
    tamd::layer_group groups[tamd::kMaxGroups];
    const uint32_t n_groups = tamd::build_layer_groups(layers, n_layers, groups);
    for (uint32_t g = 0; g < n_groups; ++g) {
        for (uint32_t k = 0; k < tamd::kGroupSize; ++k) {
            member[g * tamd::kGroupSize + k] =
                k < groups[g].n ? static_cast<int32_t>(groups[g].member[k]) : -1;
        }
        box[4 * g + 0] = groups[g].env_r[0];
        box[4 * g + 1] = groups[g].env_r[1];
        box[4 * g + 2] = groups[g].env_z[0];
        box[4 * g + 3] = groups[g].env_z[1];
        bounded[g] = groups[g].bounded;
    }
    return n_groups;
}
