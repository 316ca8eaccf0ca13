// Refit of the triangle BVH4 to new vertex positions (moving meshes): the
// triangle records are re-formed from the vertices, then every node's child
// boxes are recomputed bottom-up, one level per launch, deepest first, from
// the children's boxes (internal child) or its triangles' vertex bounds (leaf
// child) -- the bounds the host build takes, so a refit to unchanged
// positions reproduces the built boxes exactly.  The topology stays; the
// closest hit does not depend on it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dmt {

// TriRec k <- (v0, id kept), v1 - v0, v2 - v0 of the triangle at leaf slot k
// (vertex indices leaf_idx[3k .. 3k + 2]).
hipError_t launch_refit_tris(const float* positions, const uint32_t* leaf_idx, uint32_t tri_count, float4* tris,
                             hipStream_t stream);
// The child boxes of `count` nodes (indices node_list[0 .. count - 1]), whose
// internal children are already refit.
hipError_t launch_refit_nodes(float4* nodes, const uint32_t* node_list, uint32_t count, const float* positions,
                              const uint32_t* leaf_idx, hipStream_t stream);

}  // namespace dmt
