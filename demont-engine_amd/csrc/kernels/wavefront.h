// Wavefront path tracer: state, queues and launchers.
//
// Every live path carries its own bounce and sample index.  One iteration is
//   k_extend -> k_shade -> k_shadow<light> -> k_shadow<env> -> k_finish [-> k_regen]
// and k_finish turns each path that ended into the next sample of the same
// pixel, continuing its PCG state as the reference's per-pixel spp loop does
// (PathTrace.slang:5722), or into the pixel's frame value.  A pixel's samples
// therefore run strictly in order while the queue stays full until the frame
// drains.  Queue lengths live in device memory; the host never waits inside a
// frame except to stop launching once the live queue is empty.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "shade.h"

namespace dmt {

constexpr uint32_t kBlock = 256;          // threads per workgroup (4 waves)
#ifndef DMT_LDS_DEPTH
#define DMT_LDS_DEPTH 16
#endif
// LDS stack entries per thread (build-time tunable).  16 entries = 16 KB per
// workgroup: k_extend (under 64 VGPRs) then holds 8 waves per SIMD; deeper
// entries spill to a per-thread global area.
constexpr uint32_t kLdsDepth = DMT_LDS_DEPTH;
constexpr uint32_t kMaxGridBlocks = 4096; // grid-stride cap: 1M threads

// counter slots
constexpr uint32_t kCtrRays = 0;          // + parity: ray queue lengths
constexpr uint32_t kCtrFinished = 2;
constexpr uint32_t kCtrTail = 3;          // k_tail: queue entries handed out
constexpr uint32_t kCtrExtend = 4;        // + parity: k_extend rays handed out beyond the first grid pass
constexpr uint32_t kCounterSlots = 8;
// Tile order of a frame (heaviest 8x8 tiles first): per log2 bucket of a
// tile's cost on the frame set's previous frame, its tile count, then the
// next free position.  Zeroed with the counters at the start of a frame.
constexpr uint32_t kTileBuckets = 34;     // cost 0, then 2^0 .. 2^32
constexpr uint32_t kTileBinCopies = 32;   // bin copies (workgroup % copies): fewer same-address atomics
constexpr uint32_t kCtrTileBins = kCounterSlots;
constexpr uint32_t kCtrTileCursor = kCtrTileBins + kTileBinCopies * kTileBuckets;
constexpr uint32_t kCounterSlotsAll = kCtrTileCursor + kTileBuckets;
constexpr uint32_t kTileOrderBlocks = 128;   // workgroups of k_tile_order
// Tiles of a frame of n pixels in any shape: ceil(w / 8) * ceil(h / 8) <= n / 8 + 1.
__host__ __device__ inline uint32_t tile_capacity(uint32_t n) { return n / 8u + 2u; }

// Shadow records live at the slot of the path that formed them (no
// compaction); the pixel word carries these flags.
constexpr uint32_t kClampBit = 0x80000000u;
constexpr uint32_t kValidBit = 0x40000000u;

// Path state, indexed by pixel (SoA, 16 B rows).
struct PathState {
    float4* T;        // throughput.xyz, prev_lambert_brdf_pdf
    float4* L;        // radiance.xyz, flags (bit0 skip_sky_on_miss, bit1 prev_was_lambert_nee)
    float4* F;        // sample sum of the frame, then the frame value for k_accumulate
    uint32_t* rng;
    uint32_t* meta;   // bounce | sample << 8
    // Per pixel: k_tail's start time (s_memrealtime, low 32 bits) while the
    // pixel runs, its elapsed time once it has ended; null: not recorded.
    uint32_t* cost = nullptr;
    // Queue position / 64 -> 8x8 tile, heaviest tiles of the set's previous
    // frame first; null: tiles in row order.
    const uint32_t* tile_order = nullptr;
};

__host__ __device__ inline uint32_t meta_pack(uint32_t b, uint32_t s, uint32_t k) { return b | (s << 8) | (k << 16); }

// Ray queue, indexed by slot.
struct RayQueue {
    float4* o;        // origin.xyz, -
    float4* d;        // direction.xyz, pixel index bits
};

// Shadow queries of one iteration, indexed by the slot of the emitting path.
struct ShadowQueue {
    float4* o;        // origin.xyz, t_max
    float4* d;        // direction.xyz, pixel | kValidBit | kClampBit
    float4* a;        // throughput*Le (or *env colour).xyz, n_dot_l
    float4* b;        // brdf.xyz, pdf (< 0: no-trace ambient record)
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
    unsigned long long ext_rays;    // counting traversal: rays of the extend kernel
    unsigned long long tail_nodes;  // counting traversal, tail kernel (extension and shadow)
    unsigned long long tail_tris;
    unsigned long long tail_spheres;
};

struct Wavefront {
    uint32_t capacity = 0;    // pixels
    PathState ps{};
    RayQueue rq[2]{};
    uint2* hits = nullptr;    // (t bits, code) per slot
    ShadowQueue sq[4]{};      // 0: light NEE, 1: env NEE; 2, 3: sky lights (allocated for such frames only)
    uint32_t* finished = nullptr;     // per slot: pixel of a path that ended, or kMiss
    uint8_t* regen = nullptr;         // per pixel: next sample pending
    uint32_t* counters = nullptr;     // kCounterSlotsAll
    uint32_t* tile_bucket = nullptr;  // per tile: log2 bucket of its cost (tile_capacity entries)
    uint32_t* tile_order = nullptr;   // the order ps.tile_order points to (tile_capacity entries)
    uint32_t* spill = nullptr;
    uint32_t spill_depth = 0;
    uint32_t spill_threads = 0;
    DeviceStats* stats = nullptr;
    uint32_t* host_count = nullptr;   // pinned: live-queue snapshots for the stop test
};

FrameArgs make_frame_args(const float pos_fovtan[4], const float fwd_aspect[4], const float right[4],
                          const float up[4], uint32_t width, uint32_t height, uint32_t frame_index, uint32_t spp,
                          uint32_t max_bounces, uint32_t reset_accum, uint32_t mis, uint32_t ltree,
                          uint32_t refract, float clamp, float ema, uint32_t flags);

uint32_t grid_for(uint32_t n);

constexpr uint32_t kNoTail = 0xFFFFFFFFu;   // every iteration by the per-bounce kernels
#ifndef DMT_TAIL_WAVES
#define DMT_TAIL_WAVES 5
#endif
constexpr uint32_t kTailWavesPerSimd = DMT_TAIL_WAVES;   // k_tail occupancy target (and workgroups per CU)
// The extension variants (sky lights, SDF clusters) carry more live state:
// four waves per SIMD give them 128 VGPRs instead of 96, and fewer spills.
#ifndef DMT_TAIL_WAVES_EXT
#define DMT_TAIL_WAVES_EXT 4
#endif
constexpr uint32_t kTailWavesExt = DMT_TAIL_WAVES_EXT;
constexpr uint32_t tail_waves(uint32_t ext) { return ext ? kTailWavesExt : kTailWavesPerSimd; }
// Traversal steps per pass of the step-granular loops (k_tail, k_persistent)
// between the wave-level bookkeeping (ballots, refill, batched shading): three
// measured 1.5 % (1.51 M-triangle 1080p) to 4.6 % (many-light scene) faster
// than one; four (with shading at 40 ready lanes) runs the batched kernel
// 2.3 % and the many-light scene 0.7 % faster than three, the 1080p frame
// within noise (two: as three); a lane that finishes in an earlier step waits
// out the pass.
#ifndef DMT_TAIL_STEPS
#define DMT_TAIL_STEPS 4
#endif
constexpr uint32_t kTailSteps = DMT_TAIL_STEPS;
// k_tail with SDF clusters: a pass of sphere-trace steps (kSdfBurst of them)
// once this many lanes trace a cluster (0: every pass steps every lane).
// Measured on the many-light scene (512 clusters): gather 8 / burst 4 runs
// 8 % faster than stepping every lane every pass; gathers of 1, 2, 4, 16, 32
// and bursts of 2 or 8 land between -18 % and +7 %.
#ifndef DMT_SDF_GATHER
#define DMT_SDF_GATHER 8
#endif
#ifndef DMT_SDF_BURST
#define DMT_SDF_BURST 4
#endif
constexpr uint32_t kSdfGather = DMT_SDF_GATHER;
constexpr uint32_t kSdfBurst = DMT_SDF_BURST;
constexpr uint32_t kPollSlots = 4;
constexpr uint32_t kPollChunk = 4;        // iterations between live-queue snapshots

struct LaunchOptions {
    hipEvent_t* extend_events = nullptr;  // a pair around each k_extend / k_tail launch, or null
    uint8_t* event_kinds = nullptr;       // per pair: 0 k_extend, 1 k_tail
    uint32_t max_events = 0;
    uint32_t events_used = 0;             // out
    hipEvent_t accum_after = nullptr;     // the accumulation pass waits for it
    bool counting = false;                // instrumented traversal
    hipEvent_t* poll_events = nullptr;    // kPollSlots events for the stop test
    uint32_t iterations = 0;              // out: iterations launched
    uint32_t tail_after = kNoTail;        // iterations before k_tail finishes the frame
    uint32_t tail_blocks = 0;             // k_tail grid (0: grid_for(pixels))
    uint32_t extend_blocks = 0;           // k_extend grid cap (0: grid_for(pixels))
    // Camera-sample features (thin lens, shutter time, hero wavelength) or a
    // translated mesh: the frame renders in one k_tail<COUNT, true> launch.
    const FeatureArgs* features = nullptr;
};

// Render one frame into accum; the final pass that folds the frame's values
// into accum waits for opts.accum_after (the previous frame on another
// stream), keeping per-pixel frame order.
hipError_t launch_frame(const SceneView& S, const FrameArgs& F, Wavefront& W, float4* accum, hipStream_t stream,
                        LaunchOptions& opts);

// Closest-hit queries: rays (8 floats each) in device memory.
struct HitOut {
    float t;
    float n[3];
    uint32_t hit, prim, mat, user_id;
};
hipError_t launch_trace_query(const SceneView& S, const float4* rays, uint32_t n, HitOut* out, Wavefront& W,
                              hipStream_t stream);
// The same with the traversal-stack spill in caller-given device memory: a
// grid of at most max_blocks workgroups (grid-stride over the rays), so spill
// must hold max_blocks * kBlock * spill_depth entries.
hipError_t launch_trace_query_spill(const SceneView& S, const float4* rays, uint32_t n, HitOut* out, uint32_t* spill,
                                    uint32_t spill_depth, uint32_t max_blocks, hipStream_t stream);

// The denoiser inputs of a frame (PathTrace.slang:5463-5648): one closest-hit
// query per pixel along the pixel-centre camera ray (no jitter), and per
// pixel the linear depth, the reprojected motion and the primary surface's
// normal, albedo, specular F0, roughness and specular hit distance.
enum GbufferOut : uint32_t {
    kGbDepth = 0,            // 1 float
    kGbMotion = 1,           // 2
    kGbNormal = 2,           // 4
    kGbAlbedo = 3,           // 4
    kGbSpecularAlbedo = 4,   // 4
    kGbRoughness = 5,        // 1
    kGbSpecularDistance = 6, // 1
    kGbCloudTrans = 7,       // 1
    kGbOutputs = 8,
};
struct GbufferArgs {
    FrameArgs F;
    float vp_curr[16], vp_prev[16];   // column-major: clip = M * (p, 1)
    float* out[kGbOutputs];           // null: not written
    uint32_t stride;                  // floats from one pixel to the next
};
hipError_t launch_gbuffer(const SceneView& S, const GbufferArgs& A, Wavefront& W, hipStream_t stream);

}  // namespace dmt
