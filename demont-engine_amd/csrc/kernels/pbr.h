// PBR material atlas (PathTrace.slang:1435-1605, applyPbrTextures; the
// software device's CPU mirror SoftwareTracer.cpp:356-525).
//
// The engine packs every texture into one RGBA8 atlas, kPbrAtlasW texels
// wide and kPbrTile rows per tile: a tile's kPbrTile-square base image on the
// left, its mips 1..kPbrMaxLod (box-filtered in byte space) stacked in the
// right column (Engine::BuildPbrTileMips).  A texel reads as byte * (1/255),
// the software device's RGBA8 expansion.  A hit's albedo, normal, roughness
// and metallic maps are sampled trilinearly at the hit's UV (the sphere's
// longitude/latitude map, the plane's world-space planar map, the mesh's
// barycentric vertex UVs) and the mip level the ray cone's footprint selects.
//
// Included by path.h after shade.h.
#pragma once

namespace dmt {

constexpr uint32_t kPbrTile = 256u;                     // kPbrTileSize
constexpr uint32_t kPbrAtlasW = kPbrTile + kPbrTile / 2u;  // kPbrAtlasWidth (384)
constexpr uint32_t kPbrMaxLod = 8u;                     // log2(kPbrTile)
constexpr uint32_t kPbrNoTile = 0xFFFFFFFFu;            // kPbrNoTexTile

__device__ __forceinline__ float4 pbr_texel(const SceneView& S, uint32_t x, uint32_t y) {
    const uchar4 c = S.atlas[(size_t)y * kPbrAtlasW + x];
    const float k = 1.0f / 255.0f;
    return make_float4((float)c.x * k, (float)c.y * k, (float)c.z * k, (float)c.w * k);
}

__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t) {
    return make_float4(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z), a.w + t * (b.w - a.w));
}

// pbrMipRegion + samplePbrTileLevel: bilinear within mip `level` of the tile,
// the taps wrapped at the level's resolution (seamless tiling at every mip).
__device__ __forceinline__ float4 pbr_level(const SceneView& S, uint32_t tile, float u, float v, uint32_t level) {
    const uint32_t row_base = tile * kPbrTile;
    const uint32_t dim = level == 0u ? kPbrTile : (kPbrTile >> level);
    const uint32_t x0 = level == 0u ? 0u : kPbrTile;
    const uint32_t y0 = level == 0u ? row_base : row_base + (kPbrTile - 2u * dim);
    const float fu = u - floorf(u), fv = v - floorf(v);
    const float fx = fu * (float)dim - 0.5f, fy = fv * (float)dim - 0.5f;
    const float fx0 = floorf(fx), fy0 = floorf(fy);
    const float tx = fx - fx0, ty = fy - fy0;
    const int D = (int)dim;
    int ix0 = (int)fx0, iy0 = (int)fy0;
    const int ix1 = ((ix0 + 1) % D + D) % D, iy1 = ((iy0 + 1) % D + D) % D;
    ix0 = (ix0 % D + D) % D;
    iy0 = (iy0 % D + D) % D;
    const float4 c00 = pbr_texel(S, x0 + (uint32_t)ix0, y0 + (uint32_t)iy0);
    const float4 c10 = pbr_texel(S, x0 + (uint32_t)ix1, y0 + (uint32_t)iy0);
    const float4 c01 = pbr_texel(S, x0 + (uint32_t)ix0, y0 + (uint32_t)iy1);
    const float4 c11 = pbr_texel(S, x0 + (uint32_t)ix1, y0 + (uint32_t)iy1);
    return mix4(mix4(c00, c10, tx), mix4(c01, c11, tx), ty);
}

// samplePbrTileLinear: trilinear between the two levels around lod.
__device__ __forceinline__ float4 pbr_linear(const SceneView& S, uint32_t tile, float u, float v, float lod) {
    const float l = fminf(fmaxf(lod, 0.0f), (float)kPbrMaxLod);
    const uint32_t l0 = (uint32_t)floorf(l);
    const uint32_t l1 = l0 + 1u < kPbrMaxLod ? l0 + 1u : kPbrMaxLod;
    const float f = l - (float)l0;
    const float4 a = pbr_level(S, tile, u, v, l0);
    if (f <= 0.0f) return a;
    return mix4(a, pbr_level(S, tile, u, v, l1), f);
}

// srgbToLinear (IEC 61966-2-1 EOTF), pow as exp(y log x).
__device__ __forceinline__ float srgb_to_linear(float c) {
    const float lo = c / 12.92f;
    const float hi = c_pow(fmaxf((c + 0.055f) / 1.055f, 0.0f), 2.4f);
    return (c <= 0.04045f) ? lo : hi;
}

// pbrConeTexLod: the cone's width stretched by 1/|cos| at the surface, in
// texels of the base level; log2 as log(x) * log2(e).
__device__ __forceinline__ float pbr_cone_lod(float cone_width, f3 rd, f3 n, float uv_scale) {
    if (uv_scale <= 0.0f) return 0.0f;
    const float cos_i = fmaxf(fabsf(dot(rd, n)), 1e-3f);
    const float texels = ((cone_width / cos_i) * uv_scale) * (float)kPbrTile;
    return fminf(fmaxf(c_log(fmaxf(texels, 1.0f)) * 1.44269504088896341f, 0.0f), (float)kPbrMaxLod);
}

// applyNormalMap: the tangent-space sample in Duff et al.'s branchless basis
// around n.
__device__ __forceinline__ f3 pbr_normal_map(f3 n, f3 ts) {
    const float s = (n.z >= 0.0f) ? 1.0f : -1.0f;
    const float a = -1.0f / (s + n.z);
    const float b = (n.x * n.y) * a;
    const f3 t{1.0f + ((s * n.x) * n.x) * a, s * b, -(s * n.x)};
    const f3 bt{b, s + (n.y * n.y) * a, -n.y};
    return normalize((t * ts.x + bt * ts.y) + n * ts.z);
}

// The barycentrics (u, v) Moller-Trumbore forms for the triangle record at
// leaf position pos (the weights of v1 and v2).
__device__ __forceinline__ void pbr_tri_bary(const SceneView& S, uint32_t pos, f3 o, f3 d, float& u, float& v) {
    const float4* tr = S.tris + (size_t)pos * 3u;
    const f3 v0 = xyz(tr[0]), e1 = xyz(tr[1]), e2 = xyz(tr[2]);
    const f3 pv = cross(d, e2);
    const float inv_det = 1.0f / dot(e1, pv);
    const f3 tv = o - v0;
    u = dot(tv, pv) * inv_det;
    v = dot(d, cross(tv, e1)) * inv_det;
}

// The hit's map tiles, UV and UV density (testAnalyticPrim's sphere and plane
// maps, PathTrace.slang:2104-2158; the mesh's vertex UVs, :2536-2543; no maps
// on SDF clusters).  hp: o + d t.
__device__ __forceinline__ void pbr_uv(const SceneView& S, f3 o, f3 d, f3 hp, HitRec h, const Surface& sf,
                                       float time, uint32_t (&tiles)[4], float& u, float& v, float& uv_scale) {
    tiles[0] = tiles[1] = tiles[2] = tiles[3] = kPbrNoTile;
    u = 0.0f;
    v = 0.0f;
    uv_scale = 0.0f;
    if (sf.kind == 1u) {
        const float4* p = S.prims + (size_t)sf.prim * 7u;
        const float4 t6 = p[6];
        tiles[0] = __float_as_uint(t6.x);
        tiles[1] = __float_as_uint(t6.y);
        tiles[2] = __float_as_uint(t6.z);
        tiles[3] = __float_as_uint(t6.w);
        const float kPI = 3.14159265358979f;
        if (p[2].x == 0.0f) {
            const f3 sn = sf.n;
            u = 0.5f + c_atan2(sn.z, sn.x) / (2.0f * kPI);
            v = 0.5f - c_asin(fminf(fmaxf(sn.y, -1.0f), 1.0f)) / kPI;
            const float sr = fmaxf(p[0].w, 1e-4f);
            const float cos_lat = fmaxf(sqrtf(sn.x * sn.x + sn.z * sn.z), 0.05f);
            uv_scale = fmaxf(1.0f / (kPI * sr), 1.0f / (((2.0f * kPI) * sr) * cos_lat));
        } else {
            const f3 an{fabsf(sf.n.x), fabsf(sf.n.y), fabsf(sf.n.z)};
            if (an.y >= an.x && an.y >= an.z) {
                u = hp.x;
                v = hp.z;
            } else if (an.x >= an.z) {
                u = hp.z;
                v = hp.y;
            } else {
                u = hp.x;
                v = hp.y;
            }
            uv_scale = 1.0f;
        }
    } else if (sf.kind == 2u && S.mesh_uv) {
        tiles[0] = S.mesh_tex[0];
        tiles[1] = S.mesh_tex[1];
        tiles[2] = S.mesh_tex[2];
        tiles[3] = S.mesh_tex[3];
        const uint32_t pos = h.code & ~kMeshBit;
        f3 om = o;
        if (S.mesh_shift) om = o - mix3(S.mesh_prev, S.mesh_curr, time);
        float bu, bv;
        pbr_tri_bary(S, pos, om, d, bu, bv);
        const uint32_t* ix = S.mesh_tri_vtx + (size_t)pos * 3u;
        const float2 q0 = S.mesh_uv[ix[0]], q1 = S.mesh_uv[ix[1]], q2 = S.mesh_uv[ix[2]];
        const float w0 = (1.0f - bu) - bv;
        u = (q0.x * w0 + q1.x * bu) + q2.x * bv;
        v = (q0.y * w0 + q1.y * bu) + q2.y * bv;
    }
}

// applyPbrTextures at a hit: the cone's LOD, then albedo (sRGB decoded, times
// the flat tint), roughness (times the flat value, saturated), metallic
// (Lambert -> metal above 0.5) and the normal map last.  A map applies when
// its tile is below the atlas's tile count (the CPU mirror's gate;
// kPbrNoTile never is).
__device__ __forceinline__ void apply_pbr(const SceneView& S, f3 o, f3 d, f3 hp, HitRec h, Surface& sf, float time,
                                          float cone_width) {
    uint32_t tiles[4];
    float u, v, uv_scale;
    pbr_uv(S, o, d, hp, h, sf, time, tiles, u, v, uv_scale);
    const uint32_t n = S.atlas_tiles;
    if (!(tiles[0] < n || tiles[1] < n || tiles[2] < n || tiles[3] < n)) return;
    const float lod = pbr_cone_lod(cone_width, d, sf.n, uv_scale);
    if (tiles[0] < n) {
        const float4 a = pbr_linear(S, tiles[0], u, v, lod);
        sf.albedo = f3{sf.albedo.x * srgb_to_linear(a.x), sf.albedo.y * srgb_to_linear(a.y),
                       sf.albedo.z * srgb_to_linear(a.z)};
    }
    if (tiles[2] < n) sf.rough = saturate(sf.rough * pbr_linear(S, tiles[2], u, v, lod).x);
    if (tiles[3] < n) {
        const float m = pbr_linear(S, tiles[3], u, v, lod).x;
        if (m > 0.5f && sf.mat == 0u) sf.mat = 1u;
    }
    if (tiles[1] < n) {
        const float4 t = pbr_linear(S, tiles[1], u, v, lod);
        sf.n = pbr_normal_map(sf.n, f3{t.x * 2.0f - 1.0f, t.y * 2.0f - 1.0f, t.z * 2.0f - 1.0f});
    }
}

}  // namespace dmt
