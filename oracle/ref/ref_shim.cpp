// C entry points over the reference's own host builders, compiled from the
// reference sources in place (see Makefile).  TEST INFRASTRUCTURE ONLY: used
// to pin the oracle's and the product's restated builders.
#include <cstdint>
#include <cstring>
#include <span>
#include <vector>

#include "renderer/AnalyticBvh.h"
#include "renderer/LightTree.h"
#include "renderer/TriangleBvh.h"

extern "C" {

// TriangleBvh::Build (src/renderer/TriangleBvh.cpp:33-91): 8 x u32 per node.
uint32_t ref_tri_bvh(const float* pos, uint32_t vc, const uint32_t* idx, uint32_t ic, uint32_t* out_nodes,
                     uint32_t* out_perm) {
    pt::renderer::TriangleBvh b;
    b.Build(std::span<const float>(pos, (size_t)vc * 3), std::span<const uint32_t>(idx, ic));
    const auto& n = b.Nodes();
    const auto& p = b.PermutedPrimIds();
    if (!n.empty()) std::memcpy(out_nodes, n.data(), n.size() * sizeof(n[0]));
    if (!p.empty()) std::memcpy(out_perm, p.data(), p.size() * 4);
    return (uint32_t)n.size();
}

// AnalyticBvh::Build (src/renderer/AnalyticBvh.cpp:369-398) over spheres
// given as (cx, cy, cz, r): 8 x u32 per node.
uint32_t ref_analytic_bvh(const float* spheres, uint32_t n, uint32_t* out_nodes, uint32_t* out_perm) {
    std::vector<pt::renderer::BvhPrim> prims(n);
    for (uint32_t i = 0; i < n; ++i) {
        prims[i].center[0] = spheres[4 * i];
        prims[i].center[1] = spheres[4 * i + 1];
        prims[i].center[2] = spheres[4 * i + 2];
        prims[i].radius = spheres[4 * i + 3];
        prims[i].prim_id = i;
    }
    pt::renderer::AnalyticBvh b;
    b.Build(prims);
    const auto& nd = b.Nodes();
    const auto& p = b.PermutedPrimIds();
    if (!nd.empty()) std::memcpy(out_nodes, nd.data(), nd.size() * sizeof(nd[0]));
    if (!p.empty()) std::memcpy(out_perm, p.data(), p.size() * 4);
    return (uint32_t)nd.size();
}

// BuildLightTree (src/renderer/LightTree.cpp:424-455) from 80 B packed
// lights, converted to LightInput as Engine::EnsureLightTreeUploaded does
// for identity orientations (Engine.cpp:5479-5508).
uint32_t ref_light_tree(const float* lights, uint32_t n, float* out_nodes) {
    std::vector<pt::renderer::LightInput> in(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* L = lights + 20 * (size_t)i;
        auto& li = in[i];
        li.type = (uint32_t)(L[7] + 0.5f);
        for (int k = 0; k < 3; ++k) {
            li.pos[k] = L[k];
            li.intensity[k] = L[4 + k];
            li.dir[k] = L[8 + k];
            li.u_vec[k] = (li.type == 1u) ? 0.0f : L[12 + k];
        }
        li.radius = L[3];
        li.cos_outer = L[11];
        li.v_half = L[15];
    }
    pt::renderer::LightTree t;
    pt::renderer::BuildLightTree(in, t);
    if (!t.nodes.empty()) std::memcpy(out_nodes, t.nodes.data(), t.nodes.size() * sizeof(t.nodes[0]));
    return (uint32_t)t.nodes.size();
}

}  // extern "C"
