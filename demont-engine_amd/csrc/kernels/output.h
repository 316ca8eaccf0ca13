// Display output of a frame: exposure, tone mapping and the sRGB transfer
// function applied to the accumulated mean, written as RGBA32F with alpha 1
// (or, for a denoiser's colour input, the tone-mapped linear image or the
// clamped HDR mean).
// This is the output stage of the reference's software path-trace kernel
// (src/rhi_software/SoftwareTracer.cpp:995-1007: SrgbOetf(TonemapDispatch(
// avg * exposure, op)) into the slot-0 output texture), run as its own
// HBM-bound pass over the accumulation buffer: 16 B read + 16 B written per
// pixel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dmt {

// Operators of r_tonemap_op (SoftwareTracer.cpp:325-334); other values
// select ACES, as the reference's dispatch does.
enum TonemapOp : uint32_t {
    kTonemapAces = 0,
    kTonemapAgx = 1,
    kTonemapKhronosNeutral = 2,
    kTonemapReinhard = 3,
    kTonemapLinear = 4,
};

// What the pass writes: the display image, the tone-mapped linear image, or
// the clamped HDR mean.
enum TonemapMode : uint32_t {
    kModeDisplay = 0,
    kModeLinear = 1,
    kModeHdr = 2,
};

hipError_t launch_tonemap(const float4* accum, uint32_t pixels, float exposure, uint32_t op, uint32_t mode,
                          float4* out, hipStream_t stream);

}  // namespace dmt
