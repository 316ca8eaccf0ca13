// SDF clusters: the reference's sphere-traced constructive solids (boxes,
// rounded boxes, tori, capsules, spheres and planes joined by smooth union,
// subtraction and intersection; shaders/SdfPrimitives.slang,
// shaders/PathTrace.slang:2163-2310), in the build without procedural ops
// (PT_SDF_PROCEDURAL_OPS 0, SdfPrimitives.slang:56-58): a node with any other
// op evaluates to 1e30 and the analytic normal is the only normal.
//
// A cluster is the engine's 28-float4 record (Engine.cpp:5076-5190):
//   c[0] = (aabb_min.xyz, node_count bits)   c[1] = (aabb_max.xyz, material bits)
//   c[2] = (albedo.rgb, roughness)           c[3] = (ior, 0, 0, 0)
//   node i at c[4 + 3 i]: (op, shape, child_a, child_b bits), params, (center.xyz, 0)
// Nodes are in post-order (children before parents); the root is the last.
// The walkers keep the node distances in registers (d0 .. d7, picked by
// child index through a select chain) rather than in an indexed local array,
// which the compiler would place in scratch memory.
//
// Included by dmt_dev.h after its vector helpers, which this file uses.
#pragma once

namespace dmt {

constexpr uint32_t kSdfMaxNodes = 8u;     // SdfPrim::kMaxNodes (SDF_MAX_NODES)
constexpr uint32_t kSdfStride = 28u;      // float4 per cluster
constexpr uint32_t kSdfBit = 0x40000000u; // hit code of a cluster: kSdfBit | cluster index

enum : uint32_t { kSdfSphere = 0u, kSdfBox = 1u, kSdfRoundBox = 2u, kSdfTorus = 3u, kSdfCapsule = 4u, kSdfPlane = 5u };
enum : uint32_t { kSdfLeaf = 0u, kSdfUnion = 1u, kSdfSubtract = 2u, kSdfIntersect = 3u, kSdfDisplace = 4u };

__device__ __forceinline__ float sgnf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ float len2(float x, float y) { return sqrtf(x * x + y * y); }

// Per-shape distances in shape-local space (SdfPrimitives.slang:163-195).
__device__ __forceinline__ float sdf_shape_dist(f3 p, uint32_t shape, float4 q) {
    if (shape == kSdfSphere) return length(p) - q.x;
    if (shape == kSdfBox || shape == kSdfRoundBox) {
        const f3 a{fabsf(p.x) - q.x, fabsf(p.y) - q.y, fabsf(p.z) - q.z};
        const float d = length(f3{fmaxf(a.x, 0.0f), fmaxf(a.y, 0.0f), fmaxf(a.z, 0.0f)}) +
                        fminf(fmaxf(a.x, fmaxf(a.y, a.z)), 0.0f);
        return shape == kSdfBox ? d : d - q.w;
    }
    if (shape == kSdfTorus) return len2(len2(p.x, p.z) - q.x, p.y) - q.y;
    if (shape == kSdfCapsule) {
        const float y = fminf(fmaxf(p.y, -q.x), q.x);
        return length(f3{p.x, p.y - y, p.z}) - q.y;
    }
    if (shape == kSdfPlane) return dot(p, xyz(q)) + q.w;
    return 1e30f;
}

// Per-shape unit gradients (SdfPrimitives.slang:203-261).
__device__ __forceinline__ f3 sdf_shape_normal(f3 p, uint32_t shape, float4 q) {
    const f3 up{0.0f, 1.0f, 0.0f};
    if (shape == kSdfSphere) {
        const float l = length(p);
        return l > 1e-12f ? p / l : up;
    }
    if (shape == kSdfBox || shape == kSdfRoundBox) {
        const f3 a{fabsf(p.x) - q.x, fabsf(p.y) - q.y, fabsf(p.z) - q.z};
        const f3 sg{sgnf(p.x), sgnf(p.y), sgnf(p.z)};
        if (fmaxf(a.x, fmaxf(a.y, a.z)) > 0.0f) {
            const f3 m{fmaxf(a.x, 0.0f), fmaxf(a.y, 0.0f), fmaxf(a.z, 0.0f)};
            const float lm = length(m);
            if (lm < 1e-12f) return sg * up;
            return (m / lm) * sg;
        }
        f3 n{0.0f, 0.0f, 0.0f};
        if (a.x > a.y && a.x > a.z) n.x = sg.x;
        else if (a.y > a.z) n.y = sg.y;
        else n.z = sg.z;
        return n;
    }
    if (shape == kSdfTorus) {
        const float lxz = len2(p.x, p.z);
        if (lxz < 1e-12f) return f3{0.0f, sgnf(p.y), 0.0f};
        const float s = q.x / lxz;
        const f3 d = p - f3{p.x * s, 0.0f, p.z * s};
        const float ld = length(d);
        return ld > 1e-12f ? d / ld : up;
    }
    if (shape == kSdfCapsule) {
        const float y = fminf(fmaxf(p.y, -q.x), q.x);
        const f3 d = p - f3{0.0f, y, 0.0f};
        const float l = length(d);
        return l > 1e-12f ? d / l : up;
    }
    if (shape == kSdfPlane) return xyz(q);
    return up;
}

// The blend factors of the smooth ops (SdfPrimitives.slang:276-301); lerp(a, b, h) = a + (b - a) h.
__device__ __forceinline__ float sdf_h(uint32_t op, float da, float db, float k) {
    if (op == kSdfUnion) return clampf(0.5f + (0.5f * (db - da)) / k, 0.0f, 1.0f);
    if (op == kSdfSubtract) return clampf(0.5f - (0.5f * (db + da)) / k, 0.0f, 1.0f);
    return clampf(0.5f - (0.5f * (db - da)) / k, 0.0f, 1.0f);
}
__device__ __forceinline__ float sdf_combine(uint32_t op, float da, float db, float k) {
    if (op == kSdfUnion) {
        if (k <= 0.0f) return fminf(da, db);
        const float h = sdf_h(op, da, db, k);
        return (db + (da - db) * h) - (k * h) * (1.0f - h);
    }
    if (op == kSdfSubtract) {
        if (k <= 0.0f) return fmaxf(da, -db);
        const float h = sdf_h(op, da, db, k);
        return (da + (-db - da) * h) + (k * h) * (1.0f - h);
    }
    if (op == kSdfIntersect) {
        if (k <= 0.0f) return fmaxf(da, db);
        const float h = sdf_h(op, da, db, k);
        return (db + (da - db) * h) + (k * h) * (1.0f - h);
    }
    if (op == kSdfDisplace) return da + k;
    return 1e30f;
}

template <class T> __device__ __forceinline__ T sdf_pick(const T (&v)[kSdfMaxNodes], uint32_t j) {
    T r = v[0];
#pragma unroll
    for (uint32_t k = 1; k < kSdfMaxNodes; ++k) r = (j == k) ? v[k] : r;
    return r;
}

// sdfClusterDist (SdfPrimitives.slang:377-400).
__device__ __forceinline__ float sdf_dist(const float4* c, uint32_t count, f3 p) {
    float dv[kSdfMaxNodes];
#pragma unroll
    for (uint32_t i = 0; i < kSdfMaxNodes; ++i) {
        dv[i] = 1e30f;
        if (i >= count) continue;
        const float4 n0 = c[4u + 3u * i], n1 = c[5u + 3u * i];
        const uint32_t op = __float_as_uint(n0.x);
        if (op == kSdfLeaf) {
            const f3 ctr = xyz(c[6u + 3u * i]);
            dv[i] = sdf_shape_dist(p - ctr, __float_as_uint(n0.y), n1);
        } else {
            dv[i] = sdf_combine(op, sdf_pick(dv, __float_as_uint(n0.z)), sdf_pick(dv, __float_as_uint(n0.w)), n1.x);
        }
    }
    return count > 0u ? sdf_pick(dv, count - 1u) : 1e30f;
}

// sdfClusterNormal (SdfPrimitives.slang:407-491), before its final normalize.
__device__ __forceinline__ f3 sdf_grad(const float4* c, uint32_t count, f3 p) {
    float dv[kSdfMaxNodes];
    f3 gv[kSdfMaxNodes];
    const f3 up{0.0f, 1.0f, 0.0f};
#pragma unroll
    for (uint32_t i = 0; i < kSdfMaxNodes; ++i) {
        float d = 1e30f;
        f3 g = up;
        if (i < count) {
            const float4 n0 = c[4u + 3u * i], n1 = c[5u + 3u * i];
            const uint32_t op = __float_as_uint(n0.x);
            if (op == kSdfLeaf) {
                const f3 pl = p - xyz(c[6u + 3u * i]);
                const uint32_t shape = __float_as_uint(n0.y);
                d = sdf_shape_dist(pl, shape, n1);
                g = sdf_shape_normal(pl, shape, n1);
            } else if (i > 0u) {
                const uint32_t ia = __float_as_uint(n0.z), ib = __float_as_uint(n0.w);
                float da = dv[0], db = dv[0];
                f3 ga = gv[0], gb = gv[0];
#pragma unroll
                for (uint32_t j = 1; j < i; ++j) {   // the children come before node i
                    if (ia == j) { da = dv[j]; ga = gv[j]; }
                    if (ib == j) { db = dv[j]; gb = gv[j]; }
                }
                const float k = n1.x;
                if (op == kSdfUnion || op == kSdfSubtract || op == kSdfIntersect) {
                    d = sdf_combine(op, da, db, k);
                    if (k <= 0.0f) {
                        if (op == kSdfUnion) g = (da < db) ? ga : gb;
                        else if (op == kSdfSubtract) g = (da > -db) ? ga : -gb;
                        else g = (da > db) ? ga : gb;
                    } else {
                        const float h = sdf_h(op, da, db, k);
                        g = op == kSdfSubtract ? normalize(lerp3(ga, -gb, h)) : normalize(lerp3(gb, ga, h));
                    }
                } else if (op == kSdfDisplace) {
                    d = da + k;
                    g = ga;
                }
            }
        }
        dv[i] = d;
        gv[i] = g;
    }
    f3 r = up;
#pragma unroll
    for (uint32_t i = 0; i < kSdfMaxNodes; ++i)
        if (count == i + 1u) r = gv[i];
    return r;
}

// The interval sdfClusterTrace (SdfPrimitives.slang:1330-1444) marches: the
// slab interval of the cluster box (sdfClusterAabbT, :1306-1316) clipped to
// [t_min, t_max]; empty when t >= t_end.  inv is the ray's safe inverse
// direction (1 / d, |d| floored at 1e-30).
__device__ __forceinline__ void sdf_interval(const float4* c, f3 o, f3 inv, float t_min, float t_max, float& t,
                                             float& t_end) {
    const float4 h0 = c[0], h1 = c[1];
    const float t0x = (h0.x - o.x) * inv.x, t1x = (h1.x - o.x) * inv.x;
    const float t0y = (h0.y - o.y) * inv.y, t1y = (h1.y - o.y) * inv.y;
    const float t0z = (h0.z - o.z) * inv.z, t1z = (h1.z - o.z) * inv.z;
    const float t_enter = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    const float t_exit = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    t = fmaxf(t_enter, t_min);
    t_end = fminf(t_exit, t_max);
}

// One step of the trace at distance t: true for a hit there (|d| < eps),
// else t advances by max(|d|, eps).
__device__ __forceinline__ bool sdf_step(const float4* c, f3 o, f3 d, float eps, float& t) {
    const f3 p = o + d * t;
    const float ad = fabsf(sdf_dist(c, __float_as_uint(c[0].w), p));
    if (ad < eps) return true;
    t += fmaxf(ad, eps);
    return false;
}

// sdfClusterTrace without the normal, which the shading step derives at
// o + d t (the point the trace stopped at): at most min(iters, 256) steps
// over the interval.
__device__ __forceinline__ bool sdf_march(const float4* c, f3 o, f3 d, f3 inv, float t_min, float t_max,
                                          uint32_t iters, float eps, float& t_hit) {
    float t, t_end;
    sdf_interval(c, o, inv, t_min, t_max, t, t_end);
    if (t >= t_end) return false;
    const uint32_t n = iters < 256u ? iters : 256u;
    for (uint32_t i = 0; i < n; ++i) {
        const float ti = t;
        if (sdf_step(c, o, d, eps, t)) {
            t_hit = ti;
            return true;
        }
        if (t >= t_end) return false;
    }
    return false;
}

}  // namespace dmt
