// Device-side building blocks of the gfx950 path tracer.
//
// Every expression that feeds a result the path tracer reports (hit distance,
// normal, radiance) is written with an explicit operation order and the whole
// translation unit is compiled with -ffp-contract=off, so each a*b+c rounds
// twice exactly as the reference's shader expressions are written
// (shaders/PathTrace.slang, shaders/PathTraceMath.slang).  Division and sqrt
// are the correctly rounded HIP defaults.  sin/cos/acos/atan2 are evaluated
// with the published single-precision Cephes algorithms (S. L. Moshier,
// Cephes Math Library: sinf.c, asinf.c, atanf.c) so that the results do not
// depend on the device math library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dmt {

// ---------------------------------------------------------------------------
// scalars and vectors
//
// The shading code is written once over a scalar type T: float for rendering,
// and a dual number (value, derivative) for the adjoint's tangent pass.  The
// helpers below are the float versions; the dual versions live with the dual
// type.  Scalar arguments next to a vector are taken as T (non-deduced), so
// float constants mix with either.

template <class T> struct Id { using type = T; };

__device__ __forceinline__ float sqrt_(float x) { return sqrtf(x); }
__device__ __forceinline__ float fmax_(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float fmin_(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float fabs_(float x) { return fabsf(x); }
__device__ __forceinline__ float primal(float x) { return x; }

template <class T> struct V3 {
    T x, y, z;
};
using f3 = V3<float>;

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
template <class T> __device__ __forceinline__ V3<T> to(f3 v) { return V3<T>{T(v.x), T(v.y), T(v.z)}; }
template <class T> __device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __device__ __forceinline__ V3<T> operator-(V3<T> a) { return V3<T>{-a.x, -a.y, -a.z}; }
template <class T> __device__ __forceinline__ V3<T> operator*(V3<T> a, V3<T> b) { return V3<T>{a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> __device__ __forceinline__ V3<T> operator*(V3<T> a, typename Id<T>::type s) {
    return V3<T>{a.x * s, a.y * s, a.z * s};
}
template <class T> __device__ __forceinline__ V3<T> operator/(V3<T> a, typename Id<T>::type s) {
    return V3<T>{a.x / s, a.y / s, a.z / s};
}
template <class T> __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> __device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
    return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> __device__ __forceinline__ T length(V3<T> a) { return sqrt_(dot(a, a)); }
template <class T> __device__ __forceinline__ V3<T> normalize(V3<T> a) {
    T inv = 1.0f / sqrt_(dot(a, a));
    return a * inv;
}
template <class T> __device__ __forceinline__ V3<T> reflect(V3<T> i, V3<T> n) {
    T d = dot(n, i);
    return i - n * (2.0f * d);
}
template <class T> __device__ __forceinline__ V3<T> refract(V3<T> i, V3<T> n, float eta) {
    T ni = dot(n, i);
    T k = 1.0f - (eta * eta) * (1.0f - ni * ni);
    if (k < 0.0f) return V3<T>{T(0.0f), T(0.0f), T(0.0f)};
    return i * T(eta) - n * (eta * ni + sqrt_(k));
}
template <class T> __device__ __forceinline__ T maxc(V3<T> a) { return fmax_(a.x, fmax_(a.y, a.z)); }
template <class T> __device__ __forceinline__ bool any_pos(V3<T> a) { return a.x > 0.0f || a.y > 0.0f || a.z > 0.0f; }
template <class T> __device__ __forceinline__ T clampf(T x, float lo, float hi) { return fmin_(fmax_(x, T(lo)), T(hi)); }
template <class T> __device__ __forceinline__ V3<T> lerp3(V3<T> a, V3<T> b, T t) { return a + (b - a) * t; }
// The shutter-time lerp of motion blur, a (1 - t) + b t: exactly b at t = 1,
// so a sample without motion blur (t = 1) sees the current positions.
__device__ __forceinline__ f3 mix3(f3 a, f3 b, float t) {
    const float u = 1.0f - t;
    return f3{a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t};
}
// quatRotate (PathTrace.slang:2028-2031).
__device__ __forceinline__ f3 quat_rotate(float4 q, f3 v) {
    f3 u{q.x, q.y, q.z};
    f3 c = cross(u, cross(u, v) + v * q.w);
    return v + c * 2.0f;
}

// ---------------------------------------------------------------------------
// Cephes single-precision transcendental functions

__device__ __forceinline__ void cephes_sincos(float xin, float* s_out, float* c_out) {
    const float DP1 = 0.78515625f;
    const float DP2 = 2.4187564849853515625e-4f;
    const float DP3 = 3.77489497744594108e-8f;
    const float FOPI = 1.27323954473516f;
    float x = xin;
    bool neg = false;
    if (x < 0.0f) { x = -x; neg = true; }
    unsigned j = (unsigned)(x * FOPI);
    float y = (float)j;
    if (j & 1u) { j += 1u; y += 1.0f; }
    j &= 7u;
    bool flip_s = neg, flip_c = false;
    if (j > 3u) { j -= 4u; flip_s = !flip_s; flip_c = !flip_c; }
    if (j > 1u) flip_c = !flip_c;
    float r = ((x - y * DP1) - y * DP2) - y * DP3;
    float z = r * r;
    float ps = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
    float pc = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z -
               0.5f * z + 1.0f;
    bool swap = (j == 1u || j == 2u);
    float s = swap ? pc : ps;
    float c = swap ? ps : pc;
    *s_out = flip_s ? -s : s;
    *c_out = flip_c ? -c : c;
}
__device__ __forceinline__ float c_sin(float x) { float s, c; cephes_sincos(x, &s, &c); return s; }
__device__ __forceinline__ float c_cos(float x) { float s, c; cephes_sincos(x, &s, &c); return c; }
__device__ __forceinline__ float asin_kernel(float z, float x) {
    return ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z +
            1.6666752422e-1f) * z * x + x;
}
__device__ __forceinline__ float c_asin(float xin) {
    float a = fabsf(xin);
    float r;
    if (a > 0.5f) {
        float z = 0.5f * (1.0f - a);
        r = asin_kernel(z, sqrtf(z));
        r = 1.5707963267948966f - (r + r);
    } else {
        r = asin_kernel(a * a, a);
    }
    return xin < 0.0f ? -r : r;
}
__device__ __forceinline__ float c_acos(float x) {
    if (x < -0.5f) return 3.14159265358979f - 2.0f * c_asin(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * c_asin(sqrtf(0.5f * (1.0f - x)));
    return 1.5707963267948966f - c_asin(x);
}
__device__ __forceinline__ float c_atan(float xin) {
    float x = fabsf(xin);
    float y;
    if (x > 2.414213562373095f) { y = 1.5707963267948966f; x = -1.0f / x; }
    else if (x > 0.4142135623730950f) { y = 0.7853981633974483f; x = (x - 1.0f) / (x + 1.0f); }
    else { y = 0.0f; }
    float z = x * x;
    y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z *
                 x + x);
    return xin < 0.0f ? -y : y;
}
__device__ __forceinline__ float c_atan2(float y, float x) {
    int code = (x < 0.0f) ? 2 : 0;
    if (y < 0.0f) code |= 1;
    if (x == 0.0f) {
        if (code & 1) return -1.5707963267948966f;
        if (y == 0.0f) return 0.0f;
        return 1.5707963267948966f;
    }
    if (y == 0.0f) return (code & 2) ? 3.14159265358979f : 0.0f;
    float w = 0.0f;
    if (code == 2) w = 3.14159265358979f;
    else if (code == 3) w = -3.14159265358979f;
    return w + c_atan(y / x);
}

// Cephes expf: n = floor(x log2(e) + 1/2), x - n ln 2 in two parts, the
// degree-6 minimax polynomial, the exact scaling by 2^n.  Below -87 the result
// is 0 (no denormals are formed), above 88.72 it is +inf.
__device__ __forceinline__ float c_exp(float xin) {
    if (xin > 88.72283905206835f) return __int_as_float(0x7f800000);
    if (xin < -87.0f) return 0.0f;
    float z = floorf(1.44269504088896341f * xin + 0.5f);
    float x = xin - z * 0.693359375f;
    x = x - z * -2.12194440e-4f;
    const int n = (int)z;
    const float xx = x * x;
    float p = (((((1.9875691500e-4f * x + 1.3981999507e-3f) * x + 8.3334519073e-3f) * x + 4.1665795894e-2f) * x +
                1.6666665459e-1f) * x + 5.0000001201e-1f);
    p = (p * xx + x) + 1.0f;
    return p * __uint_as_float((uint32_t)(n + 127) << 23);
}

// Cephes logf: x = m 2^e with m in [sqrt(1/2), sqrt(2)), the degree-9
// minimax polynomial of log(1 + f), ln 2 in two parts.  x <= 0 gives -inf.
__device__ __forceinline__ float c_log(float xin) {
    if (!(xin > 0.0f)) return -__int_as_float(0x7f800000);
    if (isinf(xin)) return xin;
    uint32_t bits = __float_as_uint(xin);
    int e = 0;
    if (bits < 0x00800000u) {   // denormal: scale into the normal range
        xin = xin * 8388608.0f;
        bits = __float_as_uint(xin);
        e = -23;
    }
    e += (int)((bits >> 23) & 0xFFu) - 126;
    float x = __uint_as_float((bits & 0x007FFFFFu) | 0x3F000000u);   // [0.5, 1)
    if (x < 0.707106781186547524f) {
        e -= 1;
        x = (x + x) - 1.0f;
    } else {
        x = x - 1.0f;
    }
    const float z = x * x;
    float y = ((((((((7.0376836292e-2f * x - 1.1514610310e-1f) * x + 1.1676998740e-1f) * x - 1.2420140846e-1f) * x +
                   1.4249322787e-1f) * x - 1.6668057665e-1f) * x + 2.0000714765e-1f) * x - 2.4999993993e-1f) * x +
               3.3333331174e-1f) * x * z;
    const float fe = (float)e;
    y = y + -2.12194440e-4f * fe;
    y = y + -0.5f * z;
    float r = x + y;
    r = r + 0.693359375f * fe;
    return r;
}
// pow(x, y) for x >= 0 as exp(y log x) with the Cephes functions (0 at x = 0
// for y > 0).
__device__ __forceinline__ float c_pow(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : __int_as_float(0x7f800000));
    return c_exp(y * c_log(x));
}
__device__ __forceinline__ float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
// smoothstep(e0, e1, x) = t^2 (3 - 2t), t = saturate((x - e0) / (e1 - e0)).
__device__ __forceinline__ float smoothstep(float e0, float e1, float x) {
    const float t = saturate((x - e0) / (e1 - e0));
    return (t * t) * (3.0f - 2.0f * t);
}

// ---------------------------------------------------------------------------
// RNG (PathTraceMath.slang:16-24)

__device__ __forceinline__ uint32_t pcg_hash(uint32_t x) {
    x = x * 747796405u + 2891336453u;
    x = ((x >> ((x >> 28u) + 4u)) ^ x) * 277803737u;
    return (x >> 22u) ^ x;
}
__device__ __forceinline__ float randf(uint32_t& seed) {
    seed = pcg_hash(seed);
    return (float)(seed & 0x00FFFFFFu) / 16777216.0f;
}

}  // namespace dmt

#include "sdf.h"

namespace dmt {

// ---------------------------------------------------------------------------
// scene view

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr uint32_t kMeshBit = 0x80000000u;

struct SceneView {
    const float4* prims;          // 7 float4 per prim, input order
    uint32_t prim_count;
    uint32_t plane_count;
    const float4* planes;         // (n.xyz, d)
    const uint32_t* plane_ids;
    const float4* sph_nodes;      // Node4 = 8 float4
    const float4* spheres;        // (c.xyz, r), leaf order
    const uint32_t* sphere_ids;
    const float4* tri_nodes;
    const float4* tris;           // 3 float4 per triangle, leaf order
    const float4* lights;         // 5 float4 per light
    uint32_t light_count;
    uint32_t ltree_count;
    const float4* ltree;          // 4 float4 per node
    const float4* env;            // texels
    const float* env_marg;
    const float* env_cond;
    uint32_t env_w, env_h;
    float env_total;
    float env_intensity;
    uint32_t env_present;
    uint32_t mesh_shift;          // the mesh is translated (set per frame)
    const float4* sph_prev;       // (previous c.xyz, r), leaf order
    f3 mesh_prev, mesh_curr;      // its translation at shutter times 0 and 1
    const float4* moon;           // moon texture (RGBA, row-major), null: none
    uint32_t moon_w, moon_h;
    // SDF clusters (sdf.h): records in input order, a BVH4 over their boxes
    // whose leaf slots hold (cluster index bits, 0, 0, 0), and the frame's
    // sphere-trace step cap and surface distance (sdf_params.y, sdf_params_f.x)
    const float4* sdf;
    uint32_t sdf_count;
    uint32_t sdf_iters;
    const float4* sdf_nodes;
    const float4* sdf_leaf;
    float sdf_eps;
    // PBR material atlas (pbr.h): RGBA8 texels, kPbrAtlasW wide, atlas_tiles
    // tiles of kPbrTile rows (null: none); per-vertex mesh UVs (null: none)
    // and the vertex indices of each leaf slot's triangle; per frame the
    // mesh's map tiles (albedo, normal, roughness, metallic) and the ray
    // cone's spread per unit of distance (2 tan(fov_y / 2) / height)
    const uchar4* atlas;
    uint32_t atlas_tiles;
    const float2* mesh_uv;
    const uint32_t* mesh_tri_vtx;
    uint32_t mesh_tex[4];
    float cone_spread;
    // water, per frame (water.h): Beer's-law absorption rgb (1/m), the
    // surfaces' default index, the wave normal's scale, amplitude, speed and
    // time, the deep-water tint
    f3 water_absorb;
    float water_ior;
    float4 wave;
    f3 deep_color;
};

// Traversal stack: `lds_depth` entries in LDS (thread-interleaved), the rest
// in a per-thread global spill area.  The two pointers carry their address
// spaces, so a push or pop compiles to an LDS instruction (or a global one on
// the spill path) rather than a generic flat access, whose latency is that of
// vector memory even when it lands in LDS.
using lds_u32 = __attribute__((address_space(3))) uint32_t;
using global_u32 = __attribute__((address_space(1))) uint32_t;

struct Stack {
    lds_u32* lds;         // this thread's column base
    uint32_t lds_stride;  // = block size
    uint32_t lds_depth;
    global_u32* spill;    // this thread's spill area
    uint32_t spill_depth;
    __device__ __forceinline__ void put(uint32_t i, uint32_t v) {
        if (i < lds_depth) lds[i * lds_stride] = v;
        else spill[i - lds_depth] = v;
    }
    __device__ __forceinline__ uint32_t get(uint32_t i) const {
        if (i < lds_depth) return lds[i * lds_stride];
        return spill[i - lds_depth];
    }
};

// column: this thread's first LDS entry (a __shared__ array + thread index).
__device__ __forceinline__ Stack stack_at(uint32_t* column, uint32_t stride, uint32_t lds_depth, uint32_t* spill,
                                          uint32_t spill_depth) {
    return Stack{(lds_u32*)column, stride, lds_depth, (global_u32*)spill, spill_depth};
}

struct TraceCounters {
    uint32_t nodes;
    uint32_t tris;
    uint32_t spheres;
};

struct HitRec {
    float t;
    uint32_t code;   // kMiss, analytic input index, or kMeshBit | leaf position
};

// PBRT's robust slab test bound: 1 + 2*gamma(3), gamma(n) = n*eps / (1 - n*eps).
__device__ __forceinline__ constexpr float robust_hi() { return 1.0f + 2.0f * (3.0f * 5.9604645e-8f / (1.0f - 3.0f * 5.9604645e-8f)); }
// 1 / robust_hi() rounded down (0x3F7FFFFA): t * robust_lo() <= t / robust_hi()
// for t >= 0 (a product that is exactly smaller rounds no higher), so the
// entry distance shrinks at least as much as by the division, at the cost
// of one multiply instead of a correctly rounded divide.
__device__ __forceinline__ constexpr float robust_lo() { return 0.99999964f; }

// Every traversal step fetches kFetch float4 (112 B) from one address: an
// interior node's six bound rows and its children, or the primitives of a leaf
// (two triangles, or four spheres).  Lanes of a wave at interior nodes and
// lanes at leaves then wait on memory together, once per step, instead of once
// for each kind in turn.  Leaf arrays carry kLeafPad zero float4 past their end
// so that a fetch at the last primitive stays inside the allocation.
constexpr uint32_t kFetch = 7u;
constexpr uint32_t kLeafPad = 8u;
constexpr uint32_t kLeafFirst = (1u << 27) - 1u;   // leaf reference: first primitive
// trace_queue hands out rays to a wave once all its lanes are idle: partial
// refills cost an atomic and two dependent loads each and measured slower.
constexpr uint32_t kRefillLanes = 64u;

// Tests the four children of a fetched node; returns the hit mask and entry
// distances.
__device__ __forceinline__ uint32_t node4_test(const float4 (&r)[kFetch], f3 o, f3 inv, float tbest, float tn[4],
                                               uint32_t ch[4]) {
    const float4 lx = r[0], hx = r[1], ly = r[2], hy = r[3], lz = r[4], hz = r[5];
    ch[0] = __float_as_uint(r[6].x); ch[1] = __float_as_uint(r[6].y);
    ch[2] = __float_as_uint(r[6].z); ch[3] = __float_as_uint(r[6].w);
    const float lxs[4] = {lx.x, lx.y, lx.z, lx.w}, hxs[4] = {hx.x, hx.y, hx.z, hx.w};
    const float lys[4] = {ly.x, ly.y, ly.z, ly.w}, hys[4] = {hy.x, hy.y, hy.z, hy.w};
    const float lzs[4] = {lz.x, lz.y, lz.z, lz.w}, hzs[4] = {hz.x, hz.y, hz.z, hz.w};
    const float g = robust_hi();
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float t0x = (lxs[i] - o.x) * inv.x, t1x = (hxs[i] - o.x) * inv.x;
        float t0y = (lys[i] - o.y) * inv.y, t1y = (hys[i] - o.y) * inv.y;
        float t0z = (lzs[i] - o.z) * inv.z, t1z = (hzs[i] - o.z) * inv.z;
        float tmin = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
        float tmax = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
        tmax = tmax * g;
        tmin = tmin * robust_lo();   // tmin < 0 only meets max(tmin, 0) and tmin <= tbest
        // non-short-circuit: straight-line code, no branch per child
        const bool hit = (ch[i] != kEmpty) & (tmax >= fmaxf(tmin, 0.0f)) & (tmin <= tbest);
        tn[i] = hit ? tmin : __builtin_huge_valf();
        mask |= hit ? (1u << i) : 0u;
    }
    return mask;
}

__device__ __forceinline__ void cswap(float& ta, uint32_t& ca, float& tb, uint32_t& cb) {
    if (tb < ta) {
        float t = ta; ta = tb; tb = t;
        uint32_t c = ca; ca = cb; cb = c;
    }
}

// Generic BVH4 walk: visits children near-to-far, pushes the farther ones.
// A leaf's primitives start at prims + first * STRIDE; `leaf(first, count, r)`
// tests the ones the fetch `r` holds, may lower `tbest`, and returns how many
// it consumed (a leaf larger than one fetch takes further steps).
template <bool COUNT, uint32_t STRIDE, class LeafFn>
__device__ __forceinline__ void bvh4_walk(const float4* nodes, const float4* prims, f3 o, f3 inv, float& tbest,
                                          Stack& st, LeafFn&& leaf, TraceCounters& tc) {
    uint32_t sp = 0;
    uint32_t cur = 0u;   // root node (internal reference 0)
    for (;;) {
        const bool at_leaf = (cur & kLeaf) != 0u;
        const uint32_t first = cur & kLeafFirst;
        const float4* p = at_leaf ? prims + (size_t)first * STRIDE : nodes + (size_t)cur * 8u;
        float4 r[kFetch];
#pragma unroll
        for (uint32_t k = 0; k < kFetch; ++k) r[k] = p[k];
        // Keep the fetch whole and ahead of the branch: left alone, the
        // compiler sinks the parts only interior nodes read (the children,
        // the .w lanes of the bound rows) into the node branch as separate
        // loads, a second memory wait for the wave.  The empty statement
        // (no instruction) needs every register of the fetch here.
#pragma unroll
        for (uint32_t k = 0; k < kFetch; ++k) asm volatile("" ::"v"(r[k].x), "v"(r[k].y), "v"(r[k].z), "v"(r[k].w));
        if (at_leaf) {
            const uint32_t count = ((cur >> 27) & 15u) + 1u;
            const uint32_t used = leaf(first, count, r);
            if (used < count) {
                cur = kLeaf | ((count - used - 1u) << 27) | (first + used);
                continue;
            }
        } else {
            if (COUNT) tc.nodes++;
            float tn[4];
            uint32_t ch[4];
            const uint32_t m = node4_test(r, o, inv, tbest, tn, ch);
            if (m) {
                // sort (t, child) ascending; misses carry +inf and sink.
                cswap(tn[0], ch[0], tn[1], ch[1]);
                cswap(tn[2], ch[2], tn[3], ch[3]);
                cswap(tn[0], ch[0], tn[2], ch[2]);
                cswap(tn[1], ch[1], tn[3], ch[3]);
                cswap(tn[1], ch[1], tn[2], ch[2]);
                // push the farther hits, farthest first
                const uint32_t n = __popc(m);
                if (n > 3u) st.put(sp++, ch[3]);
                if (n > 2u) st.put(sp++, ch[2]);
                if (n > 1u) st.put(sp++, ch[1]);
                cur = ch[0];
                continue;
            }
        }
        if (sp == 0) break;
        cur = st.get(--sp);
    }
}

// Closest analytic hit so far: distance, input index (tie-break), hit code.
struct AnalyticBest {
    float t;
    uint32_t id;
    uint32_t code;
};

// Closest triangle hit so far, limited to the analytic distance.
struct TriBest {
    float t;
    uint32_t id;    // a triangle at exactly the analytic distance never wins
    uint32_t pos;   // leaf position, kMiss for none
};

// Planes first (intersectPlane, |n.d| >= 1e-4, t > 1e-3).
__device__ __forceinline__ AnalyticBest planes_closest(const SceneView& S, f3 o, f3 d) {
    AnalyticBest b{1e30f, 0xFFFFFFFFu, kMiss};
    for (uint32_t i = 0; i < S.plane_count; ++i) {
        const float4 p = S.planes[i];
        const uint32_t id = S.plane_ids[i];
        f3 n = xyz(p);
        float denom = dot(d, n);
        if (fabsf(denom) < 1e-4f) continue;
        float t = -(dot(o, n) + p.w) / denom;
        if (t > 1e-3f && (t < b.t || (t == b.t && id < b.id))) {
            b.t = t; b.id = id; b.code = id;
        }
    }
    return b;
}

__device__ __forceinline__ f3 safe_inverse(f3 d) {
    f3 rs{fabsf(d.x) > 1e-30f ? d.x : 1e-30f, fabsf(d.y) > 1e-30f ? d.y : 1e-30f, fabsf(d.z) > 1e-30f ? d.z : 1e-30f};
    return f3{1.0f / rs.x, 1.0f / rs.y, 1.0f / rs.z};
}

// Up to four spheres of a leaf held in the fetch (intersectSphere: closest
// root > 1e-3); returns how many it tested.  With MOTION, a ray at shutter
// time < 1 sees each centre at lerp(previous, current, time)
// (PathTrace.slang:2093).
template <bool COUNT, bool MOTION = false>
__device__ __forceinline__ uint32_t sphere_leaf(const SceneView& S, f3 o, f3 d, uint32_t first, uint32_t count,
                                                const float4 (&r)[kFetch], AnalyticBest& b, TraceCounters& tc,
                                                float time = 1.0f) {
    const uint32_t n = count < 4u ? count : 4u;
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {
        if (k >= n) break;
        const float4 s = r[k];
        if (COUNT) tc.spheres++;
        f3 c = xyz(s);
        if (MOTION && time < 1.0f) c = mix3(xyz(S.sph_prev[first + k]), c, time);
        f3 oc = o - c;
        float bb = dot(oc, d);
        float kk = dot(oc, oc) - s.w * s.w;
        float h = bb * bb - kk;
        if (h < 0.0f) continue;
        h = sqrtf(h);
        float t0 = -bb - h, t1 = -bb + h;
        float t = (t0 > 1e-3f) ? t0 : t1;
        if (t > 1e-3f && t <= b.t) {   // the id only matters for a candidate
            const uint32_t id = S.sphere_ids[first + k];
            if (t < b.t || id < b.id) {
                b.t = t; b.id = id; b.code = id;
            }
        }
    }
    return n;
}

// Up to two triangles of a leaf held in the fetch (two-sided Moller-Trumbore,
// |det| >= 1e-8, 1e-3 < t); returns how many it tested.
template <bool COUNT>
__device__ __forceinline__ uint32_t tri_leaf(f3 o, f3 d, uint32_t first, uint32_t count, const float4 (&r)[kFetch],
                                             TriBest& b, TraceCounters& tc) {
    const uint32_t n = count < 2u ? count : 2u;
#pragma unroll
    for (uint32_t k = 0; k < 2u; ++k) {
        if (k >= n) break;
        const float4 a = r[3u * k], e1v = r[3u * k + 1u], e2v = r[3u * k + 2u];
        if (COUNT) tc.tris++;
        const f3 v0 = xyz(a), e1 = xyz(e1v), e2 = xyz(e2v);
        const f3 pv = cross(d, e2);
        const float det = dot(e1, pv);
        if (fabsf(det) < 1e-8f) continue;
        const float inv_det = 1.0f / det;
        const f3 tv = o - v0;
        const float u = dot(tv, pv) * inv_det;
        if (u < 0.0f || u > 1.0f) continue;
        const f3 qv = cross(tv, e1);
        const float v = dot(d, qv) * inv_det;
        if (v < 0.0f || u + v > 1.0f) continue;
        const float t = dot(e2, qv) * inv_det;
        const uint32_t id = __float_as_uint(a.w);
        if (t > 1e-3f && (t < b.t || (t == b.t && id < b.id))) {
            b.t = t; b.id = id; b.pos = first + k;
        }
    }
    return n;
}

// How far cluster ci is traced given the closest hit b so far: to b.t, or
// just past it when b is a cluster of a higher index, which ci replaces at
// an equal distance (b.t > 1e-3 there: the next float up is the next bit
// pattern).
__device__ __forceinline__ float sdf_t_max(const AnalyticBest& b, uint32_t ci) {
    const bool tie_ok = b.code != kMiss && (b.code & kSdfBit) != 0u && (b.code & ~kSdfBit) > ci;
    return tie_ok ? __uint_as_float(__float_as_uint(b.t) + 1u) : b.t;
}

// The SDF clusters of a leaf held in the fetch (slot k: the cluster index
// bits in .x), each sphere-traced up to the closest distance so far
// (traceSdfClusters, PathTrace.slang:2212-2306, clusters after the analytic
// primitives): a cluster replaces an analytic hit only when strictly closer,
// and of two clusters at one distance the lower index wins, as in the
// reference's in-order scan.  For a later-visited lower index to reach a tie
// its trace runs to just past the current distance.  Returns how many it tested.
template <bool COUNT>
__device__ __forceinline__ uint32_t sdf_leaf(const SceneView& S, f3 o, f3 d, f3 inv, uint32_t count,
                                             const float4 (&r)[kFetch], AnalyticBest& b, TraceCounters& tc) {
    const uint32_t n = count < 4u ? count : 4u;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t ci = __float_as_uint(r[k].x);
        if (COUNT) tc.spheres++;
        float t;
        if (sdf_march(S.sdf + (size_t)ci * kSdfStride, o, d, inv, 1e-3f, sdf_t_max(b, ci), S.sdf_iters, S.sdf_eps,
                      t)) {
            b.t = t;
            b.id = 0xFFFFFFFFu;
            b.code = kSdfBit | ci;
        }
    }
    return n;
}

__device__ __forceinline__ HitRec final_hit(const AnalyticBest& a, const TriBest& t) {
    return (t.pos != kMiss) ? HitRec{t.t, kMeshBit | t.pos} : HitRec{a.t, a.code};
}

// traceScene (PathTrace.slang:2331-2693): planes, then the sphere BVH, then
// (with SDF) the SDF clusters, then the triangle BVH limited to the analytic
// hit distance.  Ties resolve to the lower primitive index; analytic hits win
// ties against clusters and triangles.  With MOTION the query is at shutter
// time `time`: moving spheres at their lerped centres, and a translated mesh
// intersected from the origin shifted by -lerp(mesh_prev, mesh_curr, time)
// (:2483-2487).
template <bool COUNT, bool MOTION = false, bool SDF = false>
__device__ __forceinline__ HitRec trace_closest(const SceneView& S, f3 o, f3 d, Stack& st, TraceCounters& tc,
                                                float time = 1.0f) {
    AnalyticBest ab = planes_closest(S, o, d);
    const f3 inv = safe_inverse(d);
    if (S.sph_nodes) {
        bvh4_walk<COUNT, 1u>(S.sph_nodes, S.spheres, o, inv, ab.t, st, [&](uint32_t first, uint32_t count,
                                                                          const float4 (&r)[kFetch]) {
            return sphere_leaf<COUNT, MOTION>(S, o, d, first, count, r, ab, tc, time);
        }, tc);
    }
    if (SDF && S.sdf_nodes) {
        bvh4_walk<COUNT, 1u>(S.sdf_nodes, S.sdf_leaf, o, inv, ab.t, st, [&](uint32_t, uint32_t count,
                                                                          const float4 (&r)[kFetch]) {
            return sdf_leaf<COUNT>(S, o, d, inv, count, r, ab, tc);
        }, tc);
    }
    TriBest tb{ab.t, 0u, kMiss};
    if (S.tri_nodes) {
        f3 om = o;
        if (MOTION && S.mesh_shift) om = o - mix3(S.mesh_prev, S.mesh_curr, time);
        bvh4_walk<COUNT, 3u>(S.tri_nodes, S.tris, om, inv, tb.t, st, [&](uint32_t first, uint32_t count,
                                                                          const float4 (&r)[kFetch]) {
            return tri_leaf<COUNT>(om, d, first, count, r, tb, tc);
        }, tc);
    }
    return final_hit(ab, tb);
}

// A closest-hit query in progress: trace_closest's walk state, kept across
// loop steps so that a lane advances its query one node or leaf at a time
// (the sphere BVH walk, then the triangle BVH walk, as one loop with a
// phase).  The tests and tie rules are trace_closest's, so the hit is too.
struct Walk {
    f3 inv;
    AnalyticBest ab;
    TriBest tb;
    uint32_t phase;   // 0 sphere BVH, 1 SDF cluster BVH, 2 triangle BVH
    uint32_t cur, sp;
    // a cluster's sphere trace in progress, one step per walk step (kMiss:
    // none): its index, the distance to test next, the end of its interval,
    // the steps left
    uint32_t sdf_ci;
    float sdf_t, sdf_end;
    uint32_t sdf_left;
};

// Sphere-trace steps a lane takes per walk step while it traces a cluster
// (the steps and their results do not depend on how they are grouped).
#ifndef DMT_SDF_STEPS
#define DMT_SDF_STEPS 1
#endif
constexpr uint32_t kSdfStepsPerWalk = DMT_SDF_STEPS;

// w.cur after an SDF leaf's last cluster: the walk continues from the stack.
constexpr uint32_t kWalkPop = 0x7FFFFFFFu;

// The hierarchy after `phase` that the scene has (3: none).  SDF: the
// query sees the SDF clusters (the feature path).
template <bool SDF>
__device__ __forceinline__ uint32_t walk_next_phase(const SceneView& S, uint32_t phase) {
    if (phase < 1u && SDF && S.sdf_nodes) return 1u;
    if (phase < 2u && S.tri_nodes) return 2u;
    return 3u;
}

// Starts the query (o, d); returns true when it is already done (no
// hierarchies: the planes decide).
template <bool SDF = false>
__device__ __forceinline__ bool walk_begin(const SceneView& S, f3 o, f3 d, Walk& w) {
    w.inv = safe_inverse(d);
    w.ab = planes_closest(S, o, d);
    w.tb = TriBest{w.ab.t, 0u, kMiss};
    w.phase = S.sph_nodes ? 0u : walk_next_phase<SDF>(S, 0u);
    w.cur = 0u;
    w.sp = 0u;
    w.sdf_ci = kMiss;
    return w.phase == 3u;
}

// After a leaf: the next stack entry, else the next hierarchy; true when the
// query is done.
template <bool SDF>
__device__ __forceinline__ bool walk_pop(const SceneView& S, Walk& w, Stack& st) {
    if (w.sp > 0u) {
        w.cur = st.get(--w.sp);
        return false;
    }
    const uint32_t next = walk_next_phase<SDF>(S, w.phase);   // this hierarchy is done
    if (next == 3u) return true;
    if (next == 2u) w.tb = TriBest{w.ab.t, 0u, kMiss};
    w.phase = next;
    w.cur = 0u;
    return false;
}

// One step of the query: a node, one leaf fetch, or one step of an SDF
// cluster's sphere trace (a lane tracing a cluster then keeps pace with the
// lanes that walk nodes).  Returns true when the query is done; its hit is
// then final_hit(w.ab, w.tb).  MOTION and SDF as in trace_closest; a query
// with SDF must have been started by walk_begin<true>.
template <bool COUNT, bool MOTION = false, bool SDF = false>
__device__ __forceinline__ bool walk_step(const SceneView& S, f3 o, f3 d, Walk& w, Stack& st, TraceCounters& tc,
                                          float time = 1.0f) {
    if (SDF && w.sdf_ci != kMiss) {
        bool done = false;
        const float4* c = S.sdf + (size_t)w.sdf_ci * kSdfStride;
        for (uint32_t k = 0; k < kSdfStepsPerWalk && !done; ++k) {
            const float t = w.sdf_t;
            if (sdf_step(c, o, d, S.sdf_eps, w.sdf_t)) {
                w.ab.t = t;
                w.ab.id = 0xFFFFFFFFu;
                w.ab.code = kSdfBit | w.sdf_ci;
                done = true;
            } else {
                done = w.sdf_t >= w.sdf_end || --w.sdf_left == 0u;
            }
        }
        if (!done) return false;
        w.sdf_ci = kMiss;
        if (w.cur != kWalkPop) return false;   // the leaf's next cluster
        return walk_pop<SDF>(S, w, st);
    }
    const bool tri = w.phase == 2u;
    const bool sdf = SDF && w.phase == 1u;
    if (MOTION && tri && S.mesh_shift) o = o - mix3(S.mesh_prev, S.mesh_curr, time);
    const bool at_leaf = (w.cur & kLeaf) != 0u;
    const uint32_t lf = w.cur & kLeafFirst;
    const float4* p = at_leaf ? (tri ? S.tris + (size_t)lf * 3u : (sdf ? S.sdf_leaf : S.spheres) + lf)
                              : (tri ? S.tri_nodes : (sdf ? S.sdf_nodes : S.sph_nodes)) + (size_t)w.cur * 8u;
    float4 r[kFetch];
#pragma unroll
    for (uint32_t k = 0; k < kFetch; ++k) r[k] = p[k];
#pragma unroll
    for (uint32_t k = 0; k < kFetch; ++k) asm volatile("" ::"v"(r[k].x), "v"(r[k].y), "v"(r[k].z), "v"(r[k].w));
    if (at_leaf) {
        const uint32_t count = ((w.cur >> 27) & 15u) + 1u;
        if (sdf) {
            // the leaf's first cluster: its interval, then its trace steps
            const uint32_t ci = __float_as_uint(r[0].x);
            if (COUNT) tc.spheres++;
            w.cur = count > 1u ? (kLeaf | ((count - 2u) << 27) | (lf + 1u)) : kWalkPop;
            float t, t_end;
            sdf_interval(S.sdf + (size_t)ci * kSdfStride, o, w.inv, 1e-3f, sdf_t_max(w.ab, ci), t, t_end);
            const uint32_t n = S.sdf_iters < 256u ? S.sdf_iters : 256u;
            if (t < t_end && n > 0u) {
                w.sdf_ci = ci;
                w.sdf_t = t;
                w.sdf_end = t_end;
                w.sdf_left = n;
                return false;
            }
            if (w.cur != kWalkPop) return false;
            return walk_pop<SDF>(S, w, st);
        }
        uint32_t used;
        if (tri) used = tri_leaf<COUNT>(o, d, lf, count, r, w.tb, tc);
        else used = sphere_leaf<COUNT, MOTION>(S, o, d, lf, count, r, w.ab, tc, time);
        if (used < count) {
            w.cur = kLeaf | ((count - used - 1u) << 27) | (lf + used);
            return false;
        }
    } else {
        if (COUNT) tc.nodes++;
        float tn[4];
        uint32_t ch[4];
        const uint32_t m = node4_test(r, o, w.inv, tri ? w.tb.t : w.ab.t, tn, ch);
        if (m) {
            cswap(tn[0], ch[0], tn[1], ch[1]);
            cswap(tn[2], ch[2], tn[3], ch[3]);
            cswap(tn[0], ch[0], tn[2], ch[2]);
            cswap(tn[1], ch[1], tn[3], ch[3]);
            cswap(tn[1], ch[1], tn[2], ch[2]);
            const uint32_t nh = __popc(m);
            if (nh > 3u) st.put(w.sp++, ch[3]);
            if (nh > 2u) st.put(w.sp++, ch[2]);
            if (nh > 1u) st.put(w.sp++, ch[1]);
            w.cur = ch[0];
            return false;
        }
    }
    return walk_pop<SDF>(S, w, st);
}

// Closest hits of the rays 0 .. n-1 of a queue, every lane on its own ray,
// rays handed out dynamically (persistent traversal with dynamic ray fetch,
// Aila and Laine 2009): each lane starts on ray `first` (its thread index;
// rays at or beyond n are none); once kRefillLanes lanes of the wave are
// idle, they take the next rays from next_base + atomicAdd(*counter, ...),
// so *counter must read 0 at launch.  A persistent grid of a few waves per
// SIMD then balances the rays' uneven lengths across the chip.
template <bool COUNT, class LoadRay, class StoreHit>
__device__ __forceinline__ void trace_queue(const SceneView& S, uint32_t first, uint32_t next_base, uint32_t n,
                                            uint32_t* counter, Stack& st, TraceCounters& tc, LoadRay&& load_ray,
                                            StoreHit&& store_hit) {
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t ray = kMiss;
    f3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 1.0f};
    Walk w{};
    auto begin = [&](uint32_t i) {
        load_ray(i, o, d);
        ray = i;
        if (walk_begin(S, o, d, w)) {
            store_hit(i, final_hit(w.ab, w.tb));
            ray = kMiss;
        }
    };
    if (first < n) begin(first);
    bool drained = next_base >= n;
    for (;;) {
        const unsigned long long idle = __ballot(ray == kMiss);
        if (!drained && (uint32_t)__popcll(idle) >= kRefillLanes) {
            const uint32_t want = (uint32_t)__popcll(idle);
            uint32_t base = 0u;
            if (idle & (1ull << lane) ? (idle & below) == 0ull : false) base = atomicAdd(counter, want);
            base = next_base + __shfl(base, __ffsll((long long)idle) - 1);
            if (ray == kMiss) {
                const uint32_t i = base + (uint32_t)__popcll(idle & below);
                if (i < n) begin(i);
            }
            if (base + want >= n) drained = true;
        }
        if (__ballot(ray != kMiss) == 0ull) {
            if (drained) break;
            continue;
        }
        if (ray == kMiss) continue;
        if (walk_step<COUNT>(S, o, d, w, st, tc)) {
            store_hit(ray, final_hit(w.ab, w.tb));
            ray = kMiss;
        }
    }
}

// Surface attributes at a hit (the HitInfo fill of testAnalyticPrim,
// PathTrace.slang:2085-2150, and of the mesh path, PathTrace.slang:2640-2668).
template <class T> struct SurfaceT {
    V3<T> n;
    f3 albedo;
    f3 emission;
    T rough;
    float ior;
    uint32_t mat;
    uint32_t user_id;
    uint32_t prim;    // analytic input index, triangle id or SDF cluster index
    uint32_t kind;    // 1 analytic, 2 mesh, 3 SDF cluster
    T t;              // hit distance
};
using Surface = SurfaceT<float>;

// The material fields of a hit (no normal): shared by the float and the dual
// surface evaluation.
template <bool SDF = false, class T>
__device__ __forceinline__ void surface_material(const SceneView& S, HitRec h, SurfaceT<T>& s) {
    if (h.code & kMeshBit) {
        const float4 a = S.tris[(size_t)(h.code & ~kMeshBit) * 3u];
        s.albedo = f3{1.0f, 0.85f, 0.45f};   // MESH_ALBEDO (PathTrace.slang:1428)
        s.emission = f3{0.0f, 0.0f, 0.0f};
        s.rough = T(0.05f);
        s.ior = 1.0f;
        s.mat = 1u;                           // MAT_METAL (PathTrace.slang:2655)
        s.user_id = 0u;
        s.prim = __float_as_uint(a.w);
        s.kind = 2u;
    } else if (SDF && (h.code & kSdfBit)) {
        // an SDF cluster hit: the cluster's material, no emission, prim id 0
        // (PathTrace.slang:2289-2303)
        const uint32_t ci = h.code & ~kSdfBit;
        const float4* c = S.sdf + (size_t)ci * kSdfStride;
        const float4 h1 = c[1], h2 = c[2], h3 = c[3];
        s.albedo = xyz(h2);
        s.rough = T(h2.w);
        s.mat = __float_as_uint(h1.w);
        s.ior = h3.x;
        s.emission = f3{0.0f, 0.0f, 0.0f};
        s.user_id = 0u;
        s.prim = ci;
        s.kind = 3u;
    } else {
        const float4* p = S.prims + (size_t)h.code * 7u;
        const float4 v1 = p[1], v2 = p[2], v4 = p[4];
        s.albedo = xyz(v1);
        s.rough = T(v1.w);
        s.mat = (uint32_t)v2.y;
        s.ior = v2.z;
        s.emission = xyz(v4);
        s.user_id = __float_as_uint(v2.w);
        s.prim = h.code;
        s.kind = 1u;
    }
}

// The normal of a sphere hit at shutter time `time` is taken about the
// centre the ray saw (PathTrace.slang:2093-2097).
// SDF: the hit may be an SDF cluster's (the feature path's traversal).
template <bool SDF = false>
__device__ __forceinline__ Surface surface_at(const SceneView& S, f3 o, f3 d, HitRec h, float time = 1.0f) {
    Surface s;
    surface_material<SDF>(S, h, s);
    s.t = h.t;
    if (h.code & kMeshBit) {
        const float4* tr = S.tris + (size_t)(h.code & ~kMeshBit) * 3u;
        s.n = normalize(cross(xyz(tr[1]), xyz(tr[2])));
    } else if (SDF && (h.code & kSdfBit)) {
        // the analytic gradient at the point the sphere trace stopped
        // (SdfPrimitives.slang:1430-1433)
        const float4* c = S.sdf + (size_t)(h.code & ~kSdfBit) * kSdfStride;
        s.n = normalize(sdf_grad(c, __float_as_uint(c[0].w), o + d * h.t));
    } else {
        const float4* p = S.prims + (size_t)h.code * 7u;
        const float4 v0 = p[0], v2 = p[2];
        if (v2.x == 0.0f) {
            f3 c = xyz(v0);
            if (time < 1.0f) c = mix3(xyz(p[3]), c, time);
            s.n = normalize((o + d * h.t) - c);
        } else {
            s.n = quat_rotate(p[5], xyz(v0));
        }
    }
    return s;
}
template <bool SDF = false>
__device__ __forceinline__ Surface surface_of(const SceneView& S, f3 o, f3 d, HitRec h, float time = 1.0f) {
    return surface_at<SDF>(S, o, d, h, time);
}

}  // namespace dmt
