// Persistent (megakernel) form of the path tracer for many scenes at once.
//
// One launch renders one frame of every scene of a batch.  The grid is sized
// to the chip (a few workgroups per CU) and loops until the work runs out:
//
// * Work items are tiles of 64 pixels (tile_w x 64/tile_w).  Each scene has a
//   tile counter in HBM; a wave takes `fetch` tiles at a time with one atomic.
// * Workgroups land on the eight XCDs round-robin (blockIdx % 8).  Scenes are
//   dealt to eight queues, one per XCD, balanced by pixel count; a wave works
//   through its own XCD's queue first, so a scene's BVH is fetched into one
//   XCD's L2, then takes what is left in the other queues.
// * A wave stays on one scene (its SceneView and FrameArgs are wave-uniform).
//   Lanes whose pixel is done take the next pixels of the wave's tiles, so
//   short and long paths share the wave; the wave switches scene only when
//   all its lanes are idle.
// * Each lane runs one path with exactly one closest-hit query per loop
//   iteration: the extension ray, or one step of a shadow ray's
//   transmittance.  Shadow queries formed at a vertex wait in per-thread
//   scratch rows and are traced in the order the wavefront kernels trace
//   them, so every pixel's arithmetic, and the image, is the same as the
//   per-scene wavefront render.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "path.h"
#include "wavefront.h"

namespace dmt {

constexpr uint32_t kXcds = 8;
constexpr uint32_t kRecRows = 4;          // float4 rows per shadow record

struct BatchScene {
    SceneView S;
    FrameArgs F;          // frame_index and reset_accum of the batch's frame 0
    FeatureArgs X;        // camera-sample features (read by the feature variant only)
    float4* accum;
    uint32_t tiles_x;     // tiles per image row
    uint32_t tiles;       // tiles of the frame (0: nothing to render)
};

struct BatchArgs {
    const BatchScene* scenes;
    const uint32_t* order;            // scene indices, grouped by XCD queue
    uint32_t queue_begin[kXcds + 1];  // queue q = order[queue_begin[q] .. queue_begin[q + 1])
    uint32_t* work;                   // per scene: tiles handed out this frame
    uint32_t* frame;                  // frames since the batch's frame 0
    uint32_t* spill;
    uint32_t spill_depth;
    uint32_t threads;                 // grid size in threads (scratch row stride)
    float4* rec;                      // kinds x kRecRows rows x threads (2 kinds, 4 with sky lights)
    DeviceStats* stats;
    uint32_t tile_w;                  // 8, 16, 32 or 64
    uint32_t tile_shift;              // log2(tile_w)
    uint32_t fetch;                   // tiles per work fetch
    uint32_t nscenes;
};

struct BatchLaunch {
    uint32_t blocks;                  // grid size in workgroups
    uint32_t stack_lds;               // LDS stack entries per thread: 16, 24 or 32
    bool features;                    // some scene's frame uses camera-sample features
    uint32_t ext;                     // kExtSky: some scene's frame has sky lights (4 record kinds
                                      // in rec); kExtSdf: some scene has SDF clusters
};

// Sets the frame counter (stream-ordered, not part of a captured frame).
hipError_t batch_set_frame(uint32_t* frame, uint32_t value, hipStream_t stream);
// One frame: reset the tile counters, the persistent kernel, advance the
// frame counter.  Kernel launches only, so the sequence can be captured
// into a graph and replayed.
hipError_t launch_batch_frame(const BatchArgs& A, const BatchLaunch& L, hipStream_t stream);

}  // namespace dmt
