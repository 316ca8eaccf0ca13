// Water (PathTrace.slang:5089-5296, 5847-5872, 6014-6050, 6768-6906, the
// build with PT_WATER_ENABLED and without the FFT ocean): a surface of
// material 3 bends its shading normal by drifting value-noise waves
// (waterShadingNormal over cloud_value_noise, PathTraceCloud.slang:31-69),
// reflects with Schlick's F0 0.02 or refracts into / out of the water
// medium; a ray in the medium loses exp(-absorb * distance) per segment and,
// escaping, returns the deep-water tint; shadow rays refract through water
// like glass and absorb along their segments in it.
//
// Included by path.h after shade.h.
#pragma once

namespace dmt {

constexpr uint32_t kMatWater = 3u;   // MAT_WATER

// cloud_hash_u32: the lowbias32 integer hash.
__device__ __forceinline__ uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// cloud_hash of a lattice point (floored coordinates): 24-bit uniform.
__device__ __forceinline__ float water_lattice(float x, float y, float z) {
    uint32_t h = lowbias32((uint32_t)(int32_t)x);
    h = lowbias32(h ^ (uint32_t)(int32_t)y);
    h = lowbias32(h ^ (uint32_t)(int32_t)z);
    return (float)(h & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float water_lerp(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float water_quintic(float t) { return ((t * t) * t) * (t * (t * 6.0f - 15.0f) + 10.0f); }

// cloud_value_noise: the quintic-smoothed trilinear blend of the eight
// lattice hashes around p.
__device__ __forceinline__ float value_noise(f3 p) {
    const float ix = floorf(p.x), iy = floorf(p.y), iz = floorf(p.z);
    const float fx = water_quintic(p.x - ix), fy = water_quintic(p.y - iy), fz = water_quintic(p.z - iz);
    const float a = water_lattice(ix, iy, iz), b = water_lattice(ix + 1.0f, iy, iz);
    const float c = water_lattice(ix, iy + 1.0f, iz), d = water_lattice(ix + 1.0f, iy + 1.0f, iz);
    const float e = water_lattice(ix, iy, iz + 1.0f), f = water_lattice(ix + 1.0f, iy, iz + 1.0f);
    const float g = water_lattice(ix, iy + 1.0f, iz + 1.0f), h = water_lattice(ix + 1.0f, iy + 1.0f, iz + 1.0f);
    return water_lerp(water_lerp(water_lerp(a, b, fx), water_lerp(c, d, fx), fy),
                      water_lerp(water_lerp(e, f, fx), water_lerp(g, h, fx), fy), fz);
}

// waterShadingNormal: the macro normal n bent against the central-difference
// gradient (step 0.15) of two octaves of value noise over the scaled xz
// plane, drifting with wave time x speed; n itself without amplitude or scale.
__device__ __forceinline__ f3 water_normal(const SceneView& S, f3 wp, f3 n) {
    const float scale = S.wave.x, amp = S.wave.y, speed = S.wave.z;
    const float t = S.wave.w * speed;
    if (amp <= 0.0f || scale <= 0.0f) return n;
    const float eps = 0.15f;
    const float px = wp.x * scale, pz = wp.z * scale;
    const float dax = 0.13f * t, daz = 0.07f * t;
    const float dbx = -0.09f * t, dbz = 0.11f * t;
    const float k = 2.02f;
    auto hgt = [&](float x, float z, float x2, float z2) {
        return value_noise(f3{x + dax, 0.0f, z + daz}) + 0.5f * value_noise(f3{x2 + dbx, 0.0f, z2 + dbz});
    };
    const float n_px = hgt(px + eps, pz, px * k + eps, pz * k);
    const float n_nx = hgt(px - eps, pz, px * k - eps, pz * k);
    const float n_pz = hgt(px, pz + eps, px * k, pz * k + eps);
    const float n_nz = hgt(px, pz - eps, px * k, pz * k - eps);
    const float dhx = (n_px - n_nx) / (2.0f * eps);
    const float dhz = (n_pz - n_nz) / (2.0f * eps);
    const f3 t1 = (fabsf(n.y) < 0.99f) ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
    const f3 tx = normalize(cross(t1, n));
    const f3 tz = cross(n, tx);
    return normalize(n - (tx * dhx + tz * dhz) * amp);
}

// Beer's law over a segment of length seg in the water.
__device__ __forceinline__ f3 water_absorb(const SceneView& S, float seg) {
    return f3{c_exp(-S.water_absorb.x * seg), c_exp(-S.water_absorb.y * seg), c_exp(-S.water_absorb.z * seg)};
}

// The index of a water surface: its own when positive, else the frame's,
// clamped to [1, 2.4].
__device__ __forceinline__ float water_ior_of(const SceneView& S, float ior) {
    return fminf(fmaxf((ior > 0.0f) ? ior : S.water_ior, 1.0f), 2.4f);
}

}  // namespace dmt
