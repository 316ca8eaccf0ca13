// On-device build of the triangle BVH4 (for meshes whose triangles change,
// where a refit does not apply): a linear BVH (Karras 2012, "Maximizing
// parallelism in the construction of BVHs, octrees, and k-d trees") over the
// 30-bit Morton codes of the triangle centroids, bottom-up boxes, then a
// level-synchronous collapse into the 4-wide nodes the traversal reads
// (each node opens its largest-area internal children until it has four;
// subtrees of at most `max_leaf` triangles become leaves).  The result has
// the layout of the host build: Node4 rows, triangle records in leaf order,
// node indices allocated level by level (so the refit's level lists are
// index ranges).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace dmt {

struct LbvhInput {
    const float* positions;   // device, 3 floats per vertex
    const uint32_t* indices;  // device, 3 per triangle
    uint32_t tri_count;       // >= 1
    float cmin[3];            // centroid bounds (host-computed)
    float cmax[3];
    uint32_t max_leaf;        // 1..8
};

struct LbvhOutput {
    float4* nodes = nullptr;        // node_count * 8 float4 (device, owned by the caller)
    uint32_t node_count = 0;
    float4* tris = nullptr;         // tri_count * 3 float4 + pad (device, owned by the caller)
    uint32_t* leaf_idx = nullptr;   // 3 per leaf slot (device, owned by the caller)
    std::vector<uint32_t> level_start;   // first node index of each level, node_count last
};

// Builds on `stream` and waits for it; allocates the outputs with hipMalloc
// (tris with `tri_pad` zero float4 after the records).  Temporaries are freed.
hipError_t build_lbvh(const LbvhInput& in, uint32_t tri_pad, LbvhOutput& out, hipStream_t stream);

}  // namespace dmt
