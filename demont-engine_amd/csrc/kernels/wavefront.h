// Wavefront state and kernel launchers of the path tracer.
//
// One frame = for each sample s: raygen -> (extend -> shade -> shadow(light)
// -> shadow(env)) x max_bounces -> sample_end.  Samples of a pixel run in
// order, so each pixel's PCG state flows from sample s to s + 1 exactly as
// the reference's per-pixel spp loop (PathTrace.slang:5722) draws it.
// All queue lengths live in device memory: a frame issues no host sync.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "shade.h"

namespace dmt {

constexpr uint32_t kBlock = 256;          // threads per workgroup (4 waves)
constexpr uint32_t kLdsDepth = 24;        // LDS stack entries per thread
constexpr uint32_t kMaxGridBlocks = 4096; // grid-stride cap: 1M threads
constexpr uint32_t kCounterSlots = 4 * 72;    // last slot: raygen slot allocator

// Path state, indexed by pixel (SoA, 16 B rows).
struct PathState {
    float4* T;        // throughput.xyz, prev_lambert_brdf_pdf
    float4* L;        // radiance.xyz, flags (bit0 skip_sky_on_miss, bit1 prev_was_lambert_nee)
    float4* F;        // frame radiance sum over samples (spp > 1)
    uint32_t* rng;
};

// Ray queue, indexed by slot.
struct RayQueue {
    float4* o;        // origin.xyz, -
    float4* d;        // direction.xyz, pixel index bits
};

// Shadow queue: one transmittance query + the factors of its contribution.
struct ShadowQueue {
    float4* o;        // origin.xyz, t_max
    float4* d;        // direction.xyz, pixel index bits
    float4* a;        // throughput*Le (or *env colour).xyz, n_dot_l
    float4* b;        // brdf.xyz, pdf
    float* w;         // MIS weight (env queue)
};

struct DeviceStats {
    unsigned long long primary;
    unsigned long long extension;
    unsigned long long shadow;
    unsigned long long nodes;       // counting traversal, shadow kernels
    unsigned long long tris;
    unsigned long long spheres;
    unsigned long long ext_nodes;   // counting traversal, extend kernel
    unsigned long long ext_tris;
    unsigned long long ext_spheres;
};

struct Wavefront {
    uint32_t capacity = 0;    // pixels
    PathState ps{};
    RayQueue rq[2]{};
    uint2* hits = nullptr;    // (t bits, code) per slot
    ShadowQueue sq[2]{};      // 0: light NEE, 1: env NEE
    uint32_t* counters = nullptr;
    uint32_t* spill = nullptr;
    uint32_t spill_depth = 0;
    uint32_t spill_threads = 0;
    DeviceStats* stats = nullptr;
};

// Counter slots: ray queue length at bounce b, light-shadow and env-shadow
// queue lengths emitted at bounce b.
__host__ __device__ inline uint32_t ctr_rays(uint32_t b) { return b; }
__host__ __device__ inline uint32_t ctr_shadow(uint32_t kind, uint32_t b) { return 72u + kind * 72u + b; }

FrameArgs make_frame_args(const float pos_fovtan[4], const float fwd_aspect[4], const float right[4],
                          const float up[4], uint32_t width, uint32_t height, uint32_t frame_index, uint32_t spp,
                          uint32_t max_bounces, uint32_t reset_accum, uint32_t mis, uint32_t ltree,
                          uint32_t refract, float clamp, float ema);

uint32_t grid_for(uint32_t n);

// Launch the kernels of one frame on `stream`.  extend_events (may be null)
// holds 2*max_bounces*spp events recorded around each extend launch.  When
// accum_after is set, the final accumulation waits for it (the previous
// frame's accumulation on another stream), keeping per-pixel frame order.
hipError_t launch_frame(const SceneView& S, const FrameArgs& F, Wavefront& W, float4* accum, hipStream_t stream,
                        hipEvent_t* extend_events, hipEvent_t accum_after);

// Closest-hit queries: rays (8 floats each) in device memory.
struct HitOut {
    float t;
    float n[3];
    uint32_t hit, prim, mat, user_id;
};
hipError_t launch_trace_query(const SceneView& S, const float4* rays, uint32_t n, HitOut* out, Wavefront& W,
                              hipStream_t stream);

// Counting re-render: same frame with the instrumented traversal.
hipError_t launch_frame_counting(const SceneView& S, const FrameArgs& F, Wavefront& W, float4* accum,
                                 hipStream_t stream);

}  // namespace dmt
