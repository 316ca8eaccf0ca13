// Host-side scene preparation for the HIP path tracer: acceleration
// structures and wire buffers in the layouts the gfx950 kernels read.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dmt {

// 4-wide BVH node, 128 B = one gfx950 L2 cache line.  Child boxes are stored
// as six float4 rows (lo.x, hi.x, lo.y, hi.y, lo.z, hi.z) so one dwordx4 load
// fetches one coordinate of all four children.
//   child[i] <  0x80000000      internal node index
//   child[i] == kEmptyChild     unused slot
//   child[i] &  0x80000000      leaf: first = bits 0..26, count - 1 = bits 27..30
struct alignas(16) Node4 {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128 B cache line");

constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafFirstMask = (1u << 27) - 1u;
inline uint32_t leaf_ref(uint32_t first, uint32_t count) { return kLeafBit | ((count - 1u) << 27) | first; }

// Triangle record in leaf order, 48 B: (v0.xyz, tri id bits), (e1.xyz, 0),
// (e2.xyz, 0) with e1 = v1 - v0, e2 = v2 - v0 computed in fp32 -- the same
// values PathTrace.slang:2598-2599 forms on the fly.
struct alignas(16) TriRec {
    float v0[3];
    uint32_t id;
    float e1[3];
    float pad1;
    float e2[3];
    float pad2;
};
static_assert(sizeof(TriRec) == 48, "TriRec is 3 float4");

struct BoundsItem {
    float lo[3], hi[3], c[3];
};

struct Bvh4 {
    std::vector<Node4> nodes;      // root = node 0 (empty vector = no geometry)
    std::vector<uint32_t> order;   // leaf order -> item index
    uint32_t depth = 0;            // levels of Node4 below the root (root = 1)
};

// Binned-SAH binary build, collapsed to 4-wide nodes.
Bvh4 build_bvh4(const std::vector<BoundsItem>& items, uint32_t bins, uint32_t max_leaf);

// LightTree.cpp builder (64 B nodes, 16 floats each).  Returns node floats.
std::vector<float> build_light_tree(const float* lights, uint32_t light_count);

// HdrImage.cpp BuildEnvCdf.  Returns the sin-weighted luminance total.
// mask (W*H, optional): pixels whose luminance the CDF leaves out.
double build_env_cdf(const float* rgb, uint32_t W, uint32_t H, std::vector<float>& marginal,
                     std::vector<float>& conditional, const uint8_t* mask = nullptr);

constexpr uint32_t kMaxEnvLights = 8;   // Engine::kMaxHdriLights
// The engine's HDRI light extraction (Engine.cpp:4369-4532): returns the
// number of lights (<= kMaxEnvLights) with their direction and pick
// probability (dir[i][3]) and rgb flux (col), and marks their pixels in mask
// (W*H) for build_env_cdf.  top_frac: r_hdri_extract_percentile (0.005).
uint32_t extract_env_lights(const float* rgb, uint32_t W, uint32_t H, double top_frac, float dir[][4],
                            float col[][4], uint8_t* mask);

// An importance CDF given by the caller instead of the one built from the
// env map (the engine's uploads).
struct EnvCdfInput {
    const float* marginal = nullptr;      // H
    const float* conditional = nullptr;   // W*H
    double total = 0.0;
};

// Everything the kernels read, host copies in device layout.
struct PreparedScene {
    // analytic primitives
    std::vector<float> prims;             // input order, 28 floats each
    uint32_t prim_count = 0;
    std::vector<float> planes;            // float4 (n.xyz, d) per plane
    std::vector<uint32_t> plane_ids;      // input index per plane
    Bvh4 sphere_bvh;
    std::vector<float> spheres;           // float4 (c.xyz, r) in leaf order
    std::vector<float> sphere_prev;       // float4 (previous c.xyz, r) in leaf order
    std::vector<uint32_t> sphere_ids;     // input index, leaf order
    // mesh
    uint32_t tri_count = 0;
    Bvh4 tri_bvh;                         // empty when the tree is built on the device
    std::vector<TriRec> tris;             // leaf order (likewise)
    float tri_cmin[3] = {0.0f, 0.0f, 0.0f};   // triangle centroid bounds
    float tri_cmax[3] = {0.0f, 0.0f, 0.0f};
    // lights
    std::vector<float> lights;            // 20 floats each
    uint32_t light_count = 0;
    std::vector<float> light_tree;        // 16 floats per node
    // env
    bool env_present = false;
    uint32_t env_w = 0, env_h = 0;
    std::vector<float> env_rgba;          // float4 texels
    std::vector<float> env_marginal, env_conditional;
    float env_total = 0.0f;
    float env_intensity = 1.0f;
    // SDF clusters
    std::vector<float> sdf;               // kSdfClusterFloats per cluster, input order
    uint32_t sdf_count = 0;
    Bvh4 sdf_bvh;                         // over the clusters' boxes, one cluster per leaf
    std::vector<float> sdf_leaf;          // float4 per leaf slot: (cluster index bits, 0, 0, 0)
};

// The engine's SDF cluster record (Engine.cpp:5076-5190): a 4-float4 header
// and kSdfNodeLimit nodes of 3 float4.
constexpr uint32_t kSdfClusterFloats = 112u;
constexpr uint32_t kSdfNodeLimit = 8u;   // SdfPrim::kMaxNodes

// Validates the clusters and builds their BVH into S.  Returns "" on success,
// else the error message; *unsupported is set for clusters that are valid
// in the reference but outside this build (a fractal root,
// SdfFractals.slang:55-62).
std::string prepare_sdf(const float* clusters, uint32_t count, PreparedScene& S, bool* unsupported);

// Validates and prepares.  Returns "" on success, else the error message.
std::string prepare_scene(const float* prims, uint32_t prim_count, const float* lights, uint32_t light_count,
                          const float* positions, uint32_t vertex_count, const uint32_t* indices,
                          uint32_t index_count, const float* env_rgb, uint32_t env_w, uint32_t env_h,
                          float env_intensity, uint32_t bins, uint32_t max_leaf, PreparedScene& out,
                          bool host_tri_bvh = true, const EnvCdfInput* env_cdf = nullptr);

}  // namespace dmt
