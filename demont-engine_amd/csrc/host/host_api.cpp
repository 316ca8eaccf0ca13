// Host-only entry points of dmt_scenegen.h that expose the scene builders.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../../include/dmt_scenegen.h"
#include "prep.h"

extern "C" {

dmt_status dmt_build_bvh4(const float* positions, uint32_t vertex_count, const uint32_t* indices,
                          uint32_t index_count, uint32_t bins, uint32_t max_leaf, void* out_nodes, void* out_order,
                          uint32_t* out_node_count, uint32_t* out_depth) {
    if (!out_node_count || (index_count && (!positions || !indices)) || index_count % 3u || max_leaf > 8u)
        return DMT_ERR_INVALID_ARG;
    const uint32_t tris = index_count / 3u;
    std::vector<dmt::BoundsItem> items(tris);
    for (uint32_t t = 0; t < tris; ++t) {
        const uint32_t i0 = indices[3 * t], i1 = indices[3 * t + 1], i2 = indices[3 * t + 2];
        if (i0 >= vertex_count || i1 >= vertex_count || i2 >= vertex_count) return DMT_ERR_INVALID_ARG;
        const float *v0 = positions + 3 * (size_t)i0, *v1 = positions + 3 * (size_t)i1, *v2 = positions + 3 * (size_t)i2;
        for (int a = 0; a < 3; ++a) {
            items[t].lo[a] = std::min(v0[a], std::min(v1[a], v2[a]));
            items[t].hi[a] = std::max(v0[a], std::max(v1[a], v2[a]));
            items[t].c[a] = (v0[a] + v1[a] + v2[a]) * (1.0f / 3.0f);
        }
    }
    dmt::Bvh4 b = dmt::build_bvh4(items, bins, max_leaf ? max_leaf : 4u);
    *out_node_count = (uint32_t)b.nodes.size();
    if (out_depth) *out_depth = b.depth;
    if (out_nodes && !b.nodes.empty()) std::memcpy(out_nodes, b.nodes.data(), b.nodes.size() * sizeof(dmt::Node4));
    if (out_order && !b.order.empty()) std::memcpy(out_order, b.order.data(), b.order.size() * 4u);
    return DMT_OK;
}

uint32_t dmt_build_light_tree(const float* lights, uint32_t n, float* out_nodes) {
    if (!lights || !out_nodes) return 0u;
    std::vector<float> t = dmt::build_light_tree(lights, n);
    if (!t.empty()) std::memcpy(out_nodes, t.data(), t.size() * sizeof(float));
    return (uint32_t)(t.size() / 16u);
}

double dmt_build_env_cdf(const float* rgb, uint32_t W, uint32_t H, float* om, float* oc) {
    if (!rgb || !om || !oc) return 0.0;
    std::vector<float> m, c;
    double total = dmt::build_env_cdf(rgb, W, H, m, c);
    std::memcpy(om, m.data(), m.size() * sizeof(float));
    std::memcpy(oc, c.data(), c.size() * sizeof(float));
    return total;
}

double dmt_build_env_cdf_masked(const float* rgb, uint32_t W, uint32_t H, const uint8_t* mask, float* om,
                                float* oc) {
    if (!rgb || !om || !oc) return 0.0;
    std::vector<float> m, c;
    double total = dmt::build_env_cdf(rgb, W, H, m, c, mask);
    std::memcpy(om, m.data(), m.size() * sizeof(float));
    std::memcpy(oc, c.data(), c.size() * sizeof(float));
    return total;
}

uint32_t dmt_env_extract_lights(const float* rgb, uint32_t W, uint32_t H, double top_frac, float dir[8][4],
                                float col[8][4], uint8_t* out_mask) {
    if (!rgb || !dir || !col || !out_mask) return 0u;
    for (int i = 0; i < 8; ++i)
        for (int a = 0; a < 4; ++a) dir[i][a] = col[i][a] = 0.0f;
    return dmt::extract_env_lights(rgb, W, H, top_frac, dir, col, out_mask);
}

}  // extern "C"
