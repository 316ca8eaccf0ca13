// Scene validation and device-layout preparation.
//
// Mirrors what the reference does between the engine's scene edits and the
// kernel dispatch: EnsurePrimitivesUploaded (planes first, finite prims in a
// BVH; Engine.cpp:4743-5002), RebuildMeshResources/TriangleBvh::Build
// (Engine.cpp:3752-3900), EnsureLightsUploaded/EnsureLightTreeUploaded
// (Engine.cpp:5235-5600) and ReloadEnvMap's CDF (Engine.cpp:4571-4605).
// Inputs the reference rejects are rejected here with a message (the
// reference logs and drops them: WriteBuffer range check SoftwareDevice.cpp:552-556,
// malformed triangle indices TriangleBvh.cpp:49).
#include "prep.h"

#include <cmath>
#include <cstring>
#include <string>

namespace dmt {
namespace {

bool finite_all(const float* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// quatRotate(q, v) = v + 2 * cross(u, cross(u, v) + w * v) (PathTrace.slang:2028-2031),
// evaluated with the same operation order the kernels would use.
void quat_rotate(const float* q, const float* v, float* out) {
    const float u[3] = {q[0], q[1], q[2]};
    const float w[3] = {v[0] * q[3], v[1] * q[3], v[2] * q[3]};
    const float c1[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const float s[3] = {c1[0] + w[0], c1[1] + w[1], c1[2] + w[2]};
    const float c2[3] = {u[1] * s[2] - u[2] * s[1], u[2] * s[0] - u[0] * s[2], u[0] * s[1] - u[1] * s[0]};
    for (int a = 0; a < 3; ++a) out[a] = v[a] + c2[a] * 2.0f;
}

}  // namespace

std::string prepare_scene(const float* prims, uint32_t prim_count, const float* lights, uint32_t light_count,
                          const float* positions, uint32_t vertex_count, const uint32_t* indices,
                          uint32_t index_count, const float* env_rgb, uint32_t env_w, uint32_t env_h,
                          float env_intensity, uint32_t bins, uint32_t max_leaf, PreparedScene& S,
                          bool host_tri_bvh, const EnvCdfInput* env_cdf) {
    if (prim_count && !prims) return "prims is null";
    if (light_count && !lights) return "lights is null";
    if (index_count % 3u) return "index_count must be a multiple of 3";
    if (index_count && (!indices || !positions)) return "mesh arrays are null";
    if (max_leaf > 8u) return "max_leaf must be <= 8";
    if ((env_w == 0) != (env_h == 0)) return "env map needs both width and height";
    if (env_w && !env_rgb) return "env_rgb is null";
    if ((uint64_t)index_count / 3u >= (1ull << 27)) return "too many triangles (max 2^27)";
    if (prim_count >= (1u << 27)) return "too many primitives (max 2^27)";

    // -- analytic primitives -------------------------------------------------
    S.prim_count = prim_count;
    S.prims.assign(prims, prims + (size_t)prim_count * 28u);
    std::vector<BoundsItem> sph;
    std::vector<uint32_t> sph_ids;
    for (uint32_t i = 0; i < prim_count; ++i) {
        const float* p = prims + (size_t)i * 28u;
        if (!finite_all(p, 24)) return "prim " + std::to_string(i) + " has a non-finite value";
        const float type = p[8], mat = p[9];
        if (type != 0.0f && type != 1.0f) return "prim " + std::to_string(i) + ": type must be 0 (sphere) or 1 (plane)";
        if (mat != 0.0f && mat != 1.0f && mat != 2.0f && mat != 3.0f)
            return "prim " + std::to_string(i) + ": material must be 0, 1, 2 or 3 (water)";
        if (type == 1.0f) {
            float n[3];
            quat_rotate(p + 20, p, n);
            S.planes.insert(S.planes.end(), {n[0], n[1], n[2], p[3]});
            S.plane_ids.push_back(i);
        } else {
            // The box holds the sphere at its previous centre (v3) and its
            // current one (v0), between which motion blur moves it
            // (PathTrace.slang:2093; Engine.cpp:4861-4898 sweeps the same
            // volume), widened by 2^-20 of the coordinates so that the
            // rounding of the shutter-time lerp stays inside.
            BoundsItem b{};
            for (int a = 0; a < 3; ++a) {
                const float lo = std::fmin(p[a], p[12 + a]), hi = std::fmax(p[a], p[12 + a]);
                const float pad = (std::fmax(std::fabs(lo), std::fabs(hi)) + p[3]) * 9.5367431640625e-7f;
                b.lo[a] = (lo - p[3]) - pad;
                b.hi[a] = (hi + p[3]) + pad;
                b.c[a] = p[a];
            }
            sph.push_back(b);
            sph_ids.push_back(i);
        }
    }
    if (!sph.empty()) {
        S.sphere_bvh = build_bvh4(sph, bins, max_leaf ? max_leaf : 4u);
        S.spheres.resize(sph.size() * 4);
        S.sphere_prev.resize(sph.size() * 4);
        S.sphere_ids.resize(sph.size());
        for (size_t k = 0; k < sph.size(); ++k) {
            uint32_t id = sph_ids[S.sphere_bvh.order[k]];
            const float* p = prims + (size_t)id * 28u;
            for (int a = 0; a < 3; ++a) {
                S.spheres[4 * k + a] = p[a];
                S.sphere_prev[4 * k + a] = p[12 + a];
            }
            S.spheres[4 * k + 3] = p[3];
            S.sphere_prev[4 * k + 3] = p[3];
            S.sphere_ids[k] = id;
        }
    }

    // -- triangle mesh -------------------------------------------------------
    const uint32_t tris = index_count / 3u;
    if (tris) {
        if (!finite_all(positions, (size_t)vertex_count * 3u)) return "positions contain a non-finite value";
        for (uint32_t k = 0; k < index_count; ++k)
            if (indices[k] >= vertex_count) return "index " + std::to_string(k) + " out of range";
        std::vector<BoundsItem> items(tris);
        for (uint32_t t = 0; t < tris; ++t) {
            const float* v0 = positions + (size_t)indices[3 * t] * 3u;
            const float* v1 = positions + (size_t)indices[3 * t + 1] * 3u;
            const float* v2 = positions + (size_t)indices[3 * t + 2] * 3u;
            BoundsItem& b = items[t];
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = std::fmin(v0[a], std::fmin(v1[a], v2[a]));
                b.hi[a] = std::fmax(v0[a], std::fmax(v1[a], v2[a]));
                b.c[a] = (v0[a] + v1[a] + v2[a]) * (1.0f / 3.0f);
                S.tri_cmin[a] = t ? std::fmin(S.tri_cmin[a], b.c[a]) : b.c[a];
                S.tri_cmax[a] = t ? std::fmax(S.tri_cmax[a], b.c[a]) : b.c[a];
            }
        }
        S.tri_count = tris;
        if (host_tri_bvh) {   // else the device builds the tree and the records
            S.tri_bvh = build_bvh4(items, bins, max_leaf ? max_leaf : 4u);
            S.tris.resize(tris);
            for (uint32_t k = 0; k < tris; ++k) {
                const uint32_t t = S.tri_bvh.order[k];
                const float* v0 = positions + (size_t)indices[3 * t] * 3u;
                const float* v1 = positions + (size_t)indices[3 * t + 1] * 3u;
                const float* v2 = positions + (size_t)indices[3 * t + 2] * 3u;
                TriRec& r = S.tris[k];
                for (int a = 0; a < 3; ++a) {
                    r.v0[a] = v0[a];
                    r.e1[a] = v1[a] - v0[a];
                    r.e2[a] = v2[a] - v0[a];
                }
                r.id = t;
                r.pad1 = r.pad2 = 0.0f;
            }
        }
    }

    // -- lights --------------------------------------------------------------
    S.light_count = light_count;
    S.lights.assign(lights, lights + (size_t)light_count * 20u);
    for (uint32_t i = 0; i < light_count; ++i) {
        const float* L = lights + (size_t)i * 20u;
        if (!finite_all(L, 20)) return "light " + std::to_string(i) + " has a non-finite value";
        const uint32_t type = (uint32_t)(L[7] + 0.5f);
        if (L[7] < 0.0f || type > 4u) return "light " + std::to_string(i) + ": type must be 0..4";
        if (type == 4u) {   // emissive triangle: vertices a (v0), b (v2), c (v3)
            const float e1[3] = {L[8] - L[0], L[9] - L[1], L[10] - L[2]};
            const float e2[3] = {L[12] - L[0], L[13] - L[1], L[14] - L[2]};
            const float cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                        cz = e1[0] * e2[1] - e1[1] * e2[0];
            if (!(cx * cx + cy * cy + cz * cz > 1e-20f)) return "light " + std::to_string(i) + ": degenerate triangle";
        }
    }
    S.light_tree = build_light_tree(S.lights.data(), light_count);

    // -- environment map -----------------------------------------------------
    if (env_w) {
        if ((uint64_t)env_w * env_h > (1ull << 26)) return "env map too large";
        if (!finite_all(env_rgb, (size_t)env_w * env_h * 3u)) return "env map has a non-finite texel";
        S.env_present = true;
        S.env_w = env_w;
        S.env_h = env_h;
        S.env_intensity = env_intensity;
        S.env_rgba.resize((size_t)env_w * env_h * 4u);
        for (size_t p = 0; p < (size_t)env_w * env_h; ++p) {
            S.env_rgba[4 * p + 0] = env_rgb[3 * p + 0];
            S.env_rgba[4 * p + 1] = env_rgb[3 * p + 1];
            S.env_rgba[4 * p + 2] = env_rgb[3 * p + 2];
            S.env_rgba[4 * p + 3] = 1.0f;
        }
        if (env_cdf && (env_cdf->marginal || env_cdf->conditional)) {
            if (!env_cdf->marginal || !env_cdf->conditional) return "env CDF: both marginal and conditional are needed";
            if (!finite_all(env_cdf->marginal, env_h) || !finite_all(env_cdf->conditional, (size_t)env_w * env_h))
                return "env CDF has a non-finite value";
            if (!(std::isfinite(env_cdf->total) && env_cdf->total >= 0.0)) return "env CDF total must be finite and >= 0";
            S.env_marginal.assign(env_cdf->marginal, env_cdf->marginal + env_h);
            S.env_conditional.assign(env_cdf->conditional, env_cdf->conditional + (size_t)env_w * env_h);
            S.env_total = float(env_cdf->total);
        } else {
            S.env_total = float(build_env_cdf(env_rgb, env_w, env_h, S.env_marginal, S.env_conditional));
        }
    } else if (env_cdf && (env_cdf->marginal || env_cdf->conditional)) {
        return "env CDF given without an env map";
    }
    return "";
}

std::string prepare_sdf(const float* clusters, uint32_t count, PreparedScene& S, bool* unsupported) {
    *unsupported = false;
    if (count && !clusters) return "sdf_clusters is null";
    if (count >= (1u << 27)) return "too many SDF clusters (max 2^27)";
    auto bits = [](float v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        return u;
    };
    std::vector<BoundsItem> items(count);
    for (uint32_t i = 0; i < count; ++i) {
        const float* c = clusters + (size_t)i * kSdfClusterFloats;
        const std::string tag = "SDF cluster " + std::to_string(i);
        const uint32_t nodes = bits(c[3]), mat = bits(c[7]);
        if (nodes > kSdfNodeLimit) return tag + ": node_count must be <= 8";
        if (mat > 3u) return tag + ": material must be 0, 1, 2 or 3 (water)";
        const float hdr[] = {c[0], c[1], c[2], c[4], c[5], c[6], c[8], c[9], c[10], c[11], c[12]};
        if (!finite_all(hdr, 11)) return tag + " has a non-finite box or material value";
        BoundsItem& b = items[i];
        for (int a = 0; a < 3; ++a) {
            if (!(c[a] <= c[4 + a])) return tag + ": aabb_min must be <= aabb_max";
            b.lo[a] = c[a];
            b.hi[a] = c[4 + a];
            b.c[a] = (c[a] + c[4 + a]) * 0.5f;
        }
        for (uint32_t k = 0; k < nodes; ++k) {
            const float* n = c + 16u + 12u * k;
            const uint32_t op = bits(n[0]), ca = bits(n[2]), cb = bits(n[3]);
            if (!finite_all(n + 4, 7)) return tag + ": node " + std::to_string(k) + " has a non-finite parameter";
            // children come before their op (post-order); ops 5.. are the
            // procedural ones, which evaluate to 1e30 and read no child
            if ((op >= 1u && op <= 4u && ca >= k) || (op >= 1u && op <= 3u && cb >= k))
                return tag + ": node " + std::to_string(k) + " names a child at or after itself";
        }
        if (nodes) {
            const float* root = c + 16u + 12u * (nodes - 1u);
            if (bits(root[0]) == 0u && bits(root[1]) >= 6u) {
                *unsupported = true;
                return tag + ": fractal SDF clusters are not supported";
            }
        }
    }
    S.sdf_count = count;
    S.sdf.assign(clusters, clusters + (size_t)count * kSdfClusterFloats);
    if (count) {
        S.sdf_bvh = build_bvh4(items, 0u, 1u);
        S.sdf_leaf.assign((size_t)count * 4u, 0.0f);
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t ci = S.sdf_bvh.order[k];
            std::memcpy(&S.sdf_leaf[4u * k], &ci, 4);
        }
    }
    return "";
}

}  // namespace dmt
