// Display output pass: see output.h.  Operator formulas follow
// src/rhi_software/SoftwareTracer.cpp:263-341 (their published sources: the
// Narkowicz 2015 ACES fit, Sobotka's AgX with Wrensch's sixth-order contrast
// fit, the Khronos PBR Neutral operator, Reinhard), with the same constants.
#include "output.h"

#include <algorithm>

namespace dmt {

namespace {

struct c3 {
    float x, y, z;
};

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

__device__ __forceinline__ float dot3(c3 a, c3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ __forceinline__ float aces(float c) {
    const float num = c * (2.51f * c + 0.03f);
    const float den = c * (2.43f * c + 0.59f) + 0.14f;
    return clamp01(num / den);
}

__device__ __forceinline__ float agx_contrast(float x) {
    const float x2 = x * x;
    const float x4 = x2 * x2;
    return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x -
           0.00232f;
}

__device__ __forceinline__ float agx_encode(float v) {
    const float lo = -12.47393f, hi = 4.026069f;
    const float ev = fminf(fmaxf(log2f(fmaxf(v, 1e-10f)), lo), hi);
    return agx_contrast((ev - lo) / (hi - lo));
}

__device__ c3 agx(c3 c) {
    const c3 in0{0.842479062253094f, 0.0784335999999992f, 0.0792237451477643f};
    const c3 in1{0.0423282422610123f, 0.878468636469772f, 0.0791661274605434f};
    const c3 in2{0.0423756549057051f, 0.0784336f, 0.879142973793104f};
    c3 v{agx_encode(dot3(in0, c)), agx_encode(dot3(in1, c)), agx_encode(dot3(in2, c))};
    // the look stage at its neutral settings, rounded as written
    const float luma = dot3(v, c3{0.2126f, 0.7152f, 0.0722f});
    v = c3{luma + (v.x - luma), luma + (v.y - luma), luma + (v.z - luma)};
    const c3 out0{1.19687900512017f, -0.0980208811401368f, -0.0990297440797205f};
    const c3 out1{-0.0528968517574562f, 1.15190312990417f, -0.0989611768448433f};
    const c3 out2{-0.0529716355144438f, -0.0980434501171241f, 1.15107367264116f};
    const c3 w{dot3(out0, v), dot3(out1, v), dot3(out2, v)};
    return c3{clamp01(powf(fmaxf(w.x, 0.0f), 2.2f)), clamp01(powf(fmaxf(w.y, 0.0f), 2.2f)),
              clamp01(powf(fmaxf(w.z, 0.0f), 2.2f))};
}

__device__ c3 khronos_neutral(c3 c) {
    const float start = 0.8f - 0.04f;
    const float desat = 0.15f;
    const float x = fminf(c.x, fminf(c.y, c.z));
    const float offset = (x < 0.08f) ? (x - 6.25f * x * x) : 0.04f;
    c = c3{c.x - offset, c.y - offset, c.z - offset};
    const float peak = fmaxf(c.x, fmaxf(c.y, c.z));
    if (peak < start) return c;
    const float d = 1.0f - start;
    const float new_peak = 1.0f - d * d / (peak + d - start);
    const float s = new_peak / peak;
    c = c3{c.x * s, c.y * s, c.z * s};
    const float g = 1.0f - 1.0f / (desat * (peak - new_peak) + 1.0f);
    // mix(c, new_peak, g) = c * (1 - g) + new_peak * g
    return c3{c.x * (1.0f - g) + new_peak * g, c.y * (1.0f - g) + new_peak * g, c.z * (1.0f - g) + new_peak * g};
}

__device__ __forceinline__ float srgb_oetf(float x) {
    return x <= 0.0031308f ? x * 12.92f : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}

__global__ __launch_bounds__(256) void k_tonemap(const float4* __restrict__ accum, uint32_t pixels, float exposure,
                                                 uint32_t op, uint32_t mode, float4* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < pixels; i += stride) {
        float4 a = accum[i];
        if (mode != kModeDisplay) {
            // hdr_src = max(mean, 0) (PathTrace.slang:7800-7801)
            a = make_float4(fmaxf(a.x, 0.0f), fmaxf(a.y, 0.0f), fmaxf(a.z, 0.0f), a.w);
            if (mode == kModeHdr) {
                out[i] = make_float4(a.x, a.y, a.z, 1.0f);
                continue;
            }
        }
        c3 c{a.x * exposure, a.y * exposure, a.z * exposure};
        switch (op) {
            case kTonemapAgx: c = agx(c); break;
            case kTonemapKhronosNeutral: c = khronos_neutral(c); break;
            case kTonemapReinhard:
                c = c3{clamp01(c.x / (1.0f + c.x)), clamp01(c.y / (1.0f + c.y)), clamp01(c.z / (1.0f + c.z))};
                break;
            case kTonemapLinear: c = c3{clamp01(c.x), clamp01(c.y), clamp01(c.z)}; break;
            default: c = c3{aces(c.x), aces(c.y), aces(c.z)}; break;
        }
        out[i] = mode == kModeLinear ? make_float4(c.x, c.y, c.z, 1.0f)
                                     : make_float4(srgb_oetf(c.x), srgb_oetf(c.y), srgb_oetf(c.z), 1.0f);
    }
}

}  // namespace

hipError_t launch_tonemap(const float4* accum, uint32_t pixels, float exposure, uint32_t op, uint32_t mode,
                          float4* out, hipStream_t stream) {
    if (pixels == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)pixels + 255u) / 256u, 1u << 16);
    hipLaunchKernelGGL(k_tonemap, dim3(blocks), dim3(256), 0, stream, accum, pixels, exposure, op, mode, out);
    return hipGetLastError();
}

}  // namespace dmt
