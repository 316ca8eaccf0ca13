// Adjoint of one frame by path replay: gradients of an image loss with
// respect to the material parameters of the analytic primitives.
//
// The caller supplies dLoss/dI for every pixel of the frame image
// I = frame_value(sum of the pixel's samples); each pixel's samples are
// re-traced from their seeds (no tape) and every radiance term's derivative
// is scattered to the parameters it depends on:
//
// * albedo (reverse mode): a term c = clampIndirect(v) depends on the albedo
//   of every Lambert / metal vertex before it (through the throughput), on
//   the albedo of its own vertex (Lambert NEE and the ambient term carry
//   albedo / pi), and on the albedo of every dielectric its shadow query
//   passes through (transmittance tint).  Scaling channel ch of v by (1 + e)
//   changes c by e * (c_ch e_ch - [c clamped, ch = argmax v] c), so each
//   term adds q = delta * c - [clamped] (delta . c) e_argmax to per-vertex
//   accumulators; a vertex's accumulator divided by its albedo is the
//   gradient (flushed when the path ends).
// * emission: a term T * Le at the hit surface, with the clampIndirect
//   Jacobian.
// * roughness: forward mode.  The replay above records, for each metal
//   vertex, the ray that reached it, its hit and the path state it arrived
//   with; a second kernel then continues each recorded vertex in dual
//   numbers from that state on, carrying d/d(roughness) through the reflected
//   ray and everything downstream of it.  The frame's pixels are processed
//   in chunks whose records fit the record list.
//
// Sampling decisions are held fixed (the derivative of the estimator with
// its random numbers fixed, away from decision boundaries); with Russian
// roulette on, its probability is treated as a constant.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "path.h"
#include "wavefront.h"

namespace dmt {

constexpr uint32_t kAdjMaxVerts = 8;       // max_bounces limit of the adjoint
constexpr uint32_t kAdjMaxDielectric = 8;  // refract_bounces limit of the adjoint
constexpr uint32_t kAdjLdsPrims = 1024;    // per-workgroup LDS accumulation up to this many prims
constexpr uint32_t kMaterialFloats = 7;    // albedo rgb, roughness, emission rgb
// A recorded metal vertex: (o, h.t), (d, h.code), (T, prev_pdf),
// (seed, prim, bounce | skip_sky << 8 | prev_lambert << 9, lambda),
// (dLoss/dI of the sample, time).
constexpr uint32_t kAdjRecFloat4 = 5;
constexpr size_t kAdjRecMaxBytes = size_t(1) << 31;   // record list limit (pixel chunks beyond)
// Register budget of the adjoint kernels: three waves per SIMD (168 VGPRs,
// a few values spilled outside the traversal loop) instead of two.
constexpr uint32_t kAdjWavesPerSimd = 3;

struct AdjointArgs {
    SceneView S;
    FrameArgs F;
    FeatureArgs X;          // camera-sample features of the frame
    uint32_t features;      // any of them on (or a translated mesh): the feature variants
    const float4* dimage;   // dLoss/dI per pixel (rgb; w unused)
    float* grads;           // prim_count x kMaterialFloats
    uint32_t prim_count;
    uint32_t metal_prims;   // analytic metal prims (0: no roughness pass)
    uint32_t* spill;
    uint32_t spill_depth;
    float4* rec;            // metal-vertex records (rec_cap x kAdjRecFloat4 float4)
    uint32_t* rec_count;    // records of the current chunk
    uint32_t rec_cap;       // records one chunk may produce at most
    uint32_t chunk_px;      // pixels per chunk (rec_cap / (spp * max_bounces))
};

hipError_t launch_adjoint(const AdjointArgs& A, hipStream_t stream);

}  // namespace dmt
