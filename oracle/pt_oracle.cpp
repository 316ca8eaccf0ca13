// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates, on the CPU, the subset of the reference path tracer that the HIP
// path implements:
//   * RNG                 shaders/PathTraceMath.slang:16-24, seed PathTrace.slang:5457
//   * camera ray          shaders/PathTrace.slang:5729-5816 (jitter 5733; hero wavelength
//                         5737-5759, shutter time 5761-5773, thin lens 5779-5816;
//                         ptConcentricDisk / ptPolygonSample / ptCatEye, heroWavelength,
//                         cauchyIOR, spectralWeightSRGB: PathTraceMath.slang:96-310)
//   * motion blur         sphere centre lerp PathTrace.slang:2093, mesh translation 2461-2487
//   * sphere / plane      shaders/PathTraceMath.slang:194-214, testAnalyticPrim PathTrace.slang:2066-2161
//   * slab test           shaders/PathTrace.slang:2321-2329 (made conservative, see box_hit)
//   * traceScene          shaders/PathTrace.slang:2331-2693 (planes, analytic BVH, tri BVH, MT)
//   * Lambert branch      shaders/PathTrace.slang:6320-6682 (light NEE, env NEE + MIS, cosine bounce)
//   * metal / dielectric  shaders/PathTrace.slang:6683-6767
//   * light tree pick     shaders/PathTrace.slang:3929-4017, light sampling 4035-4179
//   * env map             shaders/PathTrace.slang:3401-3431, 3619-3735
//   * transmittance       shaders/PathTrace.slang:5089-5196
//   * Russian roulette    shaders/PathTrace.slang:6908-6913
//   * accumulation        shaders/PathTrace.slang:7707-7781
//   * median BVH build    src/renderer/TriangleBvh.cpp:33-169, src/renderer/AnalyticBvh.cpp:369-486
//   * light tree build    src/renderer/LightTree.cpp:63-455
//   * env CDF build       src/renderer/HdrImage.cpp:169-219
//
// Deliberate, documented choices where the reference leaves freedom:
//   * closest-hit ties resolve to (t, then lower primitive index), analytic
//     primitives beat triangles at equal t (the reference's mesh pass uses a
//     strict t < h.t, PathTrace.slang:2631);
//   * the slab test is widened by 2*gamma(3) (PBRT's robust bound) so the BVH
//     result equals the exhaustive scan; orc_trace_bruteforce pins that;
//   * multi-draw constructors float2(randf, randf) draw left to right;
//   * lerp(a, b, t) = a + t*(b - a), except the shutter-time lerp, which is
//     a*(1 - t) + b*t (exact at t = 1, as the reference's comments require);
//   * exp is the Cephes expf restatement (orc_math.h).
#include "oracle.h"
#include "orc_math.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <limits>
#include <thread>
#include <type_traits>
#include <vector>

using namespace orc;

namespace {

constexpr uint32_t PRIM_SPHERE = 0u, PRIM_PLANE = 1u;
constexpr uint32_t MAT_LAMBERT = 0u, MAT_METAL = 1u, MAT_DIELECTRIC = 2u, MAT_WATER = 3u;
constexpr uint32_t LIGHT_POINT = 0u, LIGHT_SPOT = 1u, LIGHT_SPHERE = 2u, LIGHT_QUAD = 3u;
constexpr uint32_t kAnalyticBvhThreshold = 16u;   // r_analytic_bvh_threshold (Engine.cpp:690)
constexpr uint32_t kMaxLeaf = 4u;                 // TriangleBvh/AnalyticBvh kMaxLeafSize

struct BNode {
    float mn[3];
    uint32_t left_first;
    float mx[3];
    uint32_t count;
};

struct WorkItem {
    float mn[3], mx[3], c[3];
    uint32_t id;
};

// Median split over centroids along the widest centroid axis, leaves <= 4
// (TriangleBvh.cpp:93-169 / AnalyticBvh.cpp:400-486).
struct MedianBvh {
    std::vector<BNode> nodes;
    std::vector<uint32_t> perm;

    void build(std::vector<WorkItem>& w) {
        nodes.clear();
        perm.clear();
        if (w.empty()) return;
        nodes.reserve(2 * w.size());
        nodes.push_back({});
        rec(w, 0, 0, (uint32_t)w.size());
        for (auto& it : w) perm.push_back(it.id);
    }
    void rec(std::vector<WorkItem>& w, uint32_t ni, uint32_t first, uint32_t count) {
        float bn[3] = {INFINITY, INFINITY, INFINITY}, bx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < count; ++i)
            for (int a = 0; a < 3; ++a) {
                bn[a] = std::min(bn[a], w[first + i].mn[a]);
                bx[a] = std::max(bx[a], w[first + i].mx[a]);
            }
        for (int a = 0; a < 3; ++a) { nodes[ni].mn[a] = bn[a]; nodes[ni].mx[a] = bx[a]; }
        if (count <= kMaxLeaf) { nodes[ni].left_first = first; nodes[ni].count = count; return; }
        float cn[3] = {INFINITY, INFINITY, INFINITY}, cx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < count; ++i)
            for (int a = 0; a < 3; ++a) {
                float c = w[first + i].c[a];
                if (c < cn[a]) cn[a] = c;
                if (c > cx[a]) cx[a] = c;
            }
        int axis = 0;
        float widest = cx[0] - cn[0];
        for (int a = 1; a < 3; ++a)
            if (cx[a] - cn[a] > widest) { widest = cx[a] - cn[a]; axis = a; }
        if (widest > 0.0f) {
            std::nth_element(w.begin() + first, w.begin() + first + count / 2, w.begin() + first + count,
                             [axis](const WorkItem& a, const WorkItem& b) { return a.c[axis] < b.c[axis]; });
        }
        uint32_t lc = count / 2;
        uint32_t li = (uint32_t)nodes.size();
        nodes[ni].left_first = li;
        nodes[ni].count = 0;
        nodes.push_back({});
        nodes.push_back({});
        rec(w, li, first, lc);
        rec(w, li + 1, first + lc, count - lc);
    }
};

struct LightTreeNode {
    float mn[3];
    uint32_t left_first;
    float mx[3];
    uint32_t count;
    float axis[3];
    float cos_half;
    float intensity;
    float pad[3];
};
static_assert(sizeof(LightTreeNode) == 64, "64 B light tree node");

struct LeafSummary {
    float mn[3], mx[3], axis[3];
    float cos_half, intensity;
    uint32_t id;
};

// MakeLeafSummary (LightTree.cpp:63-147).
// Engine::EnsureLightTreeUploaded rotates dir and u_vec by the light's
// orientation quaternion before the build (Engine.cpp:5479-5508).
void engine_rotate(const float* q, const float* v, float* o) {
    const float qx = q[0], qy = q[1], qz = q[2], qw = q[3];
    const float cx = qy * v[2] - qz * v[1] + qw * v[0];
    const float cy = qz * v[0] - qx * v[2] + qw * v[1];
    const float cz = qx * v[1] - qy * v[0] + qw * v[2];
    o[0] = v[0] + 2.0f * (qy * cz - qz * cy);
    o[1] = v[1] + 2.0f * (qz * cx - qx * cz);
    o[2] = v[2] + 2.0f * (qx * cy - qy * cx);
}

// Emissive triangle (light type 4; an extension, not a reference light type):
// vertices v0, v2, v3 in world space, one-sided along cross(b - a, c - a).
LeafSummary make_leaf_triangle(const float* L, uint32_t id) {
    LeafSummary s{};
    s.id = id;
    for (int i = 0; i < 3; ++i) {
        s.mn[i] = std::min(L[i], std::min(L[8 + i], L[12 + i]));
        s.mx[i] = std::max(L[i], std::max(L[8 + i], L[12 + i]));
    }
    const float e1[3] = {L[8] - L[0], L[9] - L[1], L[10] - L[2]};
    const float e2[3] = {L[12] - L[0], L[13] - L[1], L[14] - L[2]};
    const float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const float len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int i = 0; i < 3; ++i) s.axis[i] = c[i] / len;
    s.cos_half = 0.0f;
    s.intensity = (0.2126f * L[4] + 0.7152f * L[5] + 0.0722f * L[6]) * (0.5f * len);
    return s;
}

LeafSummary make_leaf(const float* Lin, uint32_t id) {
    if ((uint32_t)(Lin[7] + 0.5f) == 4u) return make_leaf_triangle(Lin, id);
    float L[20];
    std::memcpy(L, Lin, sizeof(L));
    engine_rotate(Lin + 16, Lin + 8, L + 8);
    if ((uint32_t)(Lin[7] + 0.5f) != LIGHT_SPOT) engine_rotate(Lin + 16, Lin + 12, L + 12);
    LeafSummary s{};
    s.id = id;
    uint32_t type = (uint32_t)(L[7] + 0.5f);
    const float* pos = L;
    float radius = L[3];
    const float* dir = L + 8;
    const float* u = L + 12;
    float v_half = L[15];
    float cos_outer = L[11];
    if (type == LIGHT_POINT || type == LIGHT_SPOT) {
        for (int i = 0; i < 3; ++i) { s.mn[i] = pos[i] - 1e-3f; s.mx[i] = pos[i] + 1e-3f; }
    } else if (type == LIGHT_SPHERE) {
        float r = std::max(radius, 1e-3f);
        for (int i = 0; i < 3; ++i) { s.mn[i] = pos[i] - r; s.mx[i] = pos[i] + r; }
    } else {
        float vh = std::fabs(v_half);
        for (int i = 0; i < 3; ++i) {
            float e = std::fabs(u[i]) + vh;
            s.mn[i] = pos[i] - e;
            s.mx[i] = pos[i] + e;
        }
    }
    if (type == LIGHT_POINT || type == LIGHT_SPHERE) {
        s.axis[0] = 0; s.axis[1] = 1; s.axis[2] = 0; s.cos_half = -1.0f;
    } else if (type == LIGHT_SPOT) {
        s.axis[0] = dir[0]; s.axis[1] = dir[1]; s.axis[2] = dir[2]; s.cos_half = cos_outer;
    } else {
        s.axis[0] = dir[0]; s.axis[1] = dir[1]; s.axis[2] = dir[2]; s.cos_half = 0.0f;
    }
    float lum = 0.2126f * L[4] + 0.7152f * L[5] + 0.0722f * L[6];
    if (type == LIGHT_SPHERE) {
        float r = std::max(radius, 1e-3f);
        lum *= 4.0f * 3.14159265358979f * r * r;
    } else if (type == LIGHT_QUAD) {
        float ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        lum *= 4.0f * ul * std::fabs(v_half);
    }
    s.intensity = lum;
    return s;
}

// CombineCones (LightTree.cpp:171-273).
void combine_cones(const float* a, float ca, const float* b, float cb, float* out, float& oc) {
    if (ca <= -0.999f || cb <= -0.999f) {
        out[0] = a[0]; out[1] = a[1]; out[2] = a[2]; oc = -1.0f;
        return;
    }
    float d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    d = std::min(1.0f, std::max(-1.0f, d));
    float td = std::acos(d);
    float ta = std::acos(std::clamp(ca, -1.0f, 1.0f));
    float tb = std::acos(std::clamp(cb, -1.0f, 1.0f));
    if (td + tb <= ta) { out[0] = a[0]; out[1] = a[1]; out[2] = a[2]; oc = ca; return; }
    if (td + ta <= tb) { out[0] = b[0]; out[1] = b[1]; out[2] = b[2]; oc = cb; return; }
    float tc = 0.5f * (ta + tb + td);
    if (tc >= 3.14159265358979f) { out[0] = a[0]; out[1] = a[1]; out[2] = a[2]; oc = -1.0f; return; }
    oc = std::cos(tc);
    float rot = tc - ta;
    float cr = std::cos(rot), sr = std::sin(rot);
    float p[3] = {b[0] - d * a[0], b[1] - d * a[1], b[2] - d * a[2]};
    float pl = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (pl > 1e-6f) {
        float inv = 1.0f / pl;
        p[0] *= inv; p[1] *= inv; p[2] *= inv;
    } else {
        out[0] = a[0]; out[1] = a[1]; out[2] = a[2];
        if (d < 0.0f) oc = -1.0f;
        return;
    }
    for (int i = 0; i < 3; ++i) out[i] = cr * a[i] + sr * p[i];
    float al = std::sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
    if (al > 1e-6f) {
        float inv = 1.0f / al;
        out[0] *= inv; out[1] *= inv; out[2] *= inv;
    }
}

// BuildRecursive (LightTree.cpp:326-420).
void lt_rec(std::vector<LeafSummary>& leaves, std::vector<LightTreeNode>& out, uint32_t b, uint32_t e,
            uint32_t dst) {
    uint32_t n = e - b;
    if (n == 1) {
        LightTreeNode& nd = out[dst];
        const LeafSummary& s = leaves[b];
        for (int i = 0; i < 3; ++i) { nd.mn[i] = s.mn[i]; nd.mx[i] = s.mx[i]; nd.axis[i] = s.axis[i]; }
        nd.cos_half = s.cos_half;
        nd.intensity = s.intensity;
        nd.left_first = s.id;
        nd.count = 1u;
        return;
    }
    float cn[3] = {INFINITY, INFINITY, INFINITY}, cx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = b; i < e; ++i)
        for (int k = 0; k < 3; ++k) {
            float c = 0.5f * (leaves[i].mn[k] + leaves[i].mx[k]);
            cn[k] = std::min(cn[k], c);
            cx[k] = std::max(cx[k], c);
        }
    int axis = 0;
    float best = cx[0] - cn[0];
    for (int k = 1; k < 3; ++k)
        if (cx[k] - cn[k] > best) { best = cx[k] - cn[k]; axis = k; }
    uint32_t mid = b + n / 2;
    if (!(best < 1e-6f)) {
        std::nth_element(leaves.begin() + b, leaves.begin() + mid, leaves.begin() + e,
                         [axis](const LeafSummary& x, const LeafSummary& y) {
                             float ca = 0.5f * (x.mn[axis] + x.mx[axis]);
                             float cb = 0.5f * (y.mn[axis] + y.mx[axis]);
                             return ca < cb;
                         });
    }
    uint32_t li = (uint32_t)out.size();
    out.emplace_back();
    out.emplace_back();
    lt_rec(leaves, out, b, mid, li);
    lt_rec(leaves, out, mid, e, li + 1);
    LightTreeNode& self = out[dst];
    const LightTreeNode& l = out[li];
    const LightTreeNode& r = out[li + 1];
    for (int i = 0; i < 3; ++i) {
        self.mn[i] = std::min(l.mn[i], r.mn[i]);
        self.mx[i] = std::max(l.mx[i], r.mx[i]);
    }
    self.intensity = l.intensity + r.intensity;
    combine_cones(l.axis, l.cos_half, r.axis, r.cos_half, self.axis, self.cos_half);
    self.left_first = li;
    self.count = 0u;
}

std::vector<LightTreeNode> build_light_tree(const float* lights, uint32_t n) {
    std::vector<LightTreeNode> out;
    if (n == 0) return out;
    std::vector<LeafSummary> leaves;
    for (uint32_t i = 0; i < n; ++i) leaves.push_back(make_leaf(lights + 20 * i, i));
    out.reserve(2 * n);
    out.emplace_back();
    lt_rec(leaves, out, 0, n, 0);
    return out;
}

// BuildEnvCdf (HdrImage.cpp:169-219).
double build_env_cdf(const float* rgb, uint32_t W, uint32_t H, std::vector<float>& marg,
                     std::vector<float>& cond) {
    marg.assign(H, 0.0f);
    cond.assign((size_t)W * H, 0.0f);
    double total = 0.0;
    for (uint32_t v = 0; v < H; ++v) {
        const double st = std::sin(3.14159265358979323846 * (double(v) + 0.5) / double(H));
        double row = 0.0;
        for (uint32_t u = 0; u < W; ++u) {
            size_t p = (size_t)v * W + u;
            float lum = 0.2126f * rgb[3 * p] + 0.7152f * rgb[3 * p + 1] + 0.0722f * rgb[3 * p + 2];
            row += double(lum) * st;
            cond[p] = float(row);
        }
        double nrm = row > 0.0 ? 1.0 / row : 0.0;
        for (uint32_t u = 0; u < W; ++u) {
            size_t p = (size_t)v * W + u;
            cond[p] = float(double(cond[p]) * nrm);
        }
        marg[v] = float(row);
        total += row;
    }
    double nrm = total > 0.0 ? 1.0 / total : 0.0;
    double pre = 0.0;
    for (uint32_t v = 0; v < H; ++v) {
        pre += marg[v];
        marg[v] = float(pre * nrm);
    }
    if (H > 0 && total > 0.0) marg[H - 1] = 1.0f;
    return total;
}

}  // namespace

struct orc_scene {
    std::vector<float> prims;
    uint32_t prim_count = 0;
    std::vector<uint32_t> planes;     // infinite prims, input order
    std::vector<uint32_t> finite;     // finite prims, input order
    MedianBvh abvh;                   // over finite prims (perm holds input indices)
    bool use_abvh = false;
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    uint32_t tri_count = 0;
    MedianBvh tbvh;
    std::vector<float> lights;
    uint32_t light_count = 0;
    std::vector<LightTreeNode> ltree;
    std::vector<float> env;          // RGB
    uint32_t W = 0, H = 0;
    bool env_present = false;
    float env_intensity = 1.0f;
    std::vector<float> marg, cond;
    float env_total = 0.0f;
    std::vector<float> moon;         // RGBA moon texture (empty: none)
    uint32_t moon_w = 0, moon_h = 0;
    std::vector<float> sdf;          // SDF clusters, 112 floats each
    uint32_t sdf_count = 0;
    std::vector<float> atlas;        // PBR atlas, RGBA floats (byte * (1/255)), 384 wide
    uint32_t atlas_w = 0, atlas_h = 0, atlas_tiles = 0;
    std::vector<float> uvs;          // mesh UVs, 2 per vertex
    // tangent direction of a PT<Dual> run: material parameter k (albedo rgb,
    // roughness, emission rgb) of analytic prim tangent_prim
    uint32_t tangent_prim = 0xFFFFFFFFu;
    uint32_t tangent_param = 0u;
};

namespace {

template <class R> struct Hit {
    bool hit = false;
    R t = R(1e30);
    V3<R> n{R(0), R(1), R(0)};
    uint32_t mat = MAT_LAMBERT;
    V3<R> albedo{R(0.5), R(0.5), R(0.5)};
    R rough = R(0), ior = R(1);
    V3<R> emission{R(0), R(0), R(0)};
    uint32_t user_id = 0;
    uint32_t kind = 0;   // 0 miss 1 analytic 2 mesh 3 SDF cluster
    uint32_t prim = 0;
};

struct Ctr {
    uint64_t primary = 0, extension = 0, shadow = 0, nodes = 0, tests = 0;
};

template <class R> inline V3<R> ld3(const float* p) { return V3<R>{R(p[0]), R(p[1]), R(p[2])}; }

// Slab test (PathTrace.slang:2321-2329) widened by PBRT's 1 + 2*gamma(3) so
// that no box containing the closest primitive hit is rejected by rounding.
template <class R> inline bool box_hit(V3<R> ro, V3<R> inv, const float* mn, const float* mx, R tbest) {
    R t0x = (R(mn[0]) - ro.x) * inv.x, t1x = (R(mx[0]) - ro.x) * inv.x;
    R t0y = (R(mn[1]) - ro.y) * inv.y, t1y = (R(mx[1]) - ro.y) * inv.y;
    R t0z = (R(mn[2]) - ro.z) * inv.z, t1z = (R(mx[2]) - ro.z) * inv.z;
    R tmin = std::max(std::max(std::min(t0x, t1x), std::min(t0y, t1y)), std::min(t0z, t1z));
    R tmax = std::min(std::min(std::max(t0x, t1x), std::max(t0y, t1y)), std::max(t0z, t1z));
    const R g = R(1) + R(2) * R(3.0 * 5.9604644775390625e-8 / (1.0 - 3.0 * 5.9604644775390625e-8));
    tmax = tmax * g;
    tmin = tmin / g;
    return tmax >= std::max(tmin, R(0)) && tmin <= tbest;
}

template <class R> inline bool better(R t, uint32_t id, const Hit<R>& h, uint32_t kind) {
    if (t < h.t) return true;
    return t == h.t && h.kind == kind && id < h.prim;
}

// testAnalyticPrim (PathTrace.slang:2066-2161).  The sphere centre is
// lerp(prev_pos, pos, shutter time) (:2093); without motion blur the time is 1
// and the centre is v0.xyz.
template <class R> void test_prim(const orc_scene& S, uint32_t i, V3<R> ro, V3<R> rd, Hit<R>& h, Ctr& c, R time) {
    const float* p = &S.prims[(size_t)i * 28];
    c.tests++;
    uint32_t type = (uint32_t)p[8];
    R t;
    if (type == PRIM_SPHERE) {
        V3<R> ce = (time < R(1)) ? mix3(ld3<R>(p + 12), ld3<R>(p), time) : ld3<R>(p);
        R rad = R(p[3]);
        V3<R> oc = ro - ce;
        R b = dot(oc, rd);
        R k = dot(oc, oc) - rad * rad;
        R hh = b * b - k;
        if (hh < R(0)) return;
        hh = osqrt(hh);
        R t0 = -b - hh, t1 = -b + hh;
        t = (t0 > R(1e-3)) ? t0 : t1;
        if (!(t > R(1e-3))) return;
        if (!better(t, i, h, 1u)) return;
        h.hit = true; h.t = t;
        h.n = normalize((ro + rd * t) - ce);
    } else if (type == PRIM_PLANE) {
        R q[4] = {R(p[20]), R(p[21]), R(p[22]), R(p[23])};
        V3<R> pn = quat_rotate(q, ld3<R>(p));
        R denom = dot(rd, pn);
        if (oabs(denom) < R(1e-4)) return;
        t = -(dot(ro, pn) + R(p[3])) / denom;
        if (!(t > R(1e-3))) return;
        if (!better(t, i, h, 1u)) return;
        h.hit = true; h.t = t;
        h.n = pn;
    } else {
        return;
    }
    h.mat = (uint32_t)p[9];
    h.albedo = ld3<R>(p + 4);
    h.rough = R(p[7]);
    h.ior = R(p[10]);
    h.emission = ld3<R>(p + 16);
    h.user_id = f2u(p[11]);
    h.kind = 1u;
    h.prim = i;
    if constexpr (std::is_same_v<R, Dual>) {
        if (i == S.tangent_prim) {
            switch (S.tangent_param) {
                case 0: h.albedo.x.d = 1.0; break;
                case 1: h.albedo.y.d = 1.0; break;
                case 2: h.albedo.z.d = 1.0; break;
                case 3: h.rough.d = 1.0; break;
                case 4: h.emission.x.d = 1.0; break;
                case 5: h.emission.y.d = 1.0; break;
                default: h.emission.z.d = 1.0; break;
            }
        }
    }
}

template <class R> struct TriT {
    bool ok;
    R t;
};

// Moller-Trumbore as in PathTrace.slang:2600-2625.
template <class R> inline TriT<R> tri_test(const orc_scene& S, uint32_t tri, V3<R> ro, V3<R> rd) {
    const uint32_t* ix = &S.idx[(size_t)tri * 3];
    V3<R> v0 = ld3<R>(&S.pos[(size_t)ix[0] * 3]);
    V3<R> v1 = ld3<R>(&S.pos[(size_t)ix[1] * 3]);
    V3<R> v2 = ld3<R>(&S.pos[(size_t)ix[2] * 3]);
    V3<R> e1 = v1 - v0, e2 = v2 - v0;
    V3<R> pv = cross(rd, e2);
    R det = dot(e1, pv);
    if (oabs(det) < R(1e-8)) return {false, R(0)};
    R inv_det = R(1) / det;
    V3<R> tv = ro - v0;
    R u = dot(tv, pv) * inv_det;
    if (u < R(0) || u > R(1)) return {false, R(0)};
    V3<R> qv = cross(tv, e1);
    R v = dot(rd, qv) * inv_det;
    if (v < R(0) || u + v > R(1)) return {false, R(0)};
    R t = dot(e2, qv) * inv_det;
    return {t > R(1e-3), t};
}

// The barycentrics (u, v) Moller-Trumbore forms for triangle tri and the ray
// (the weights of v1 and v2 in the mesh UV interpolation).
template <class R> inline void tri_bary(const orc_scene& S, uint32_t tri, V3<R> ro, V3<R> rd, R& u, R& v) {
    const uint32_t* ix = &S.idx[(size_t)tri * 3];
    V3<R> v0 = ld3<R>(&S.pos[(size_t)ix[0] * 3]);
    V3<R> v1 = ld3<R>(&S.pos[(size_t)ix[1] * 3]);
    V3<R> v2 = ld3<R>(&S.pos[(size_t)ix[2] * 3]);
    V3<R> e1 = v1 - v0, e2 = v2 - v0;
    V3<R> pv = cross(rd, e2);
    R inv_det = R(1) / dot(e1, pv);
    V3<R> tv = ro - v0;
    u = dot(tv, pv) * inv_det;
    v = dot(rd, cross(tv, e1)) * inv_det;
}

template <class R> void mesh_fill(const orc_scene& S, uint32_t tri, R t, Hit<R>& h) {
    const uint32_t* ix = &S.idx[(size_t)tri * 3];
    V3<R> v0 = ld3<R>(&S.pos[(size_t)ix[0] * 3]);
    V3<R> v1 = ld3<R>(&S.pos[(size_t)ix[1] * 3]);
    V3<R> v2 = ld3<R>(&S.pos[(size_t)ix[2] * 3]);
    h.hit = true;
    h.t = t;
    h.n = normalize(cross(v1 - v0, v2 - v0));
    h.mat = MAT_METAL;                               // PathTrace.slang:2655-2658
    h.albedo = V3<R>{R(1.0f), R(0.85f), R(0.45f)};   // MESH_ALBEDO (PathTrace.slang:1428)
    h.rough = R(0.05f);
    h.ior = R(1);
    h.emission = V3<R>{R(0), R(0), R(0)};
    h.user_id = 0u;
    h.kind = 2u;
    h.prim = tri;
}

template <class R> inline V3<R> safe_inv(V3<R> rd) {
    auto f = [](R x) { return oabs(x) > R(1e-30) ? x : R(1e-30); };
    return V3<R>{R(1) / f(rd.x), R(1) / f(rd.y), R(1) / f(rd.z)};
}

// SDF clusters (shaders/SdfPrimitives.slang, traceSdfClusters
// PathTrace.slang:2212-2306) as the reference builds them without procedural
// ops (PT_SDF_PROCEDURAL_OPS 0, SdfPrimitives.slang:56-58): a node with an op
// beyond displace evaluates to 1e30, and the normal is the analytic one.
// min/max are IEEE minNum/maxNum (a NaN operand yields the other), as the
// GPU's.  Cluster record: hdr0 (aabb_min, node_count bits), hdr1 (aabb_max,
// material bits), hdr2 (albedo, roughness), hdr3 (ior), then node i at float
// 16 + 12 i: (op, shape, child_a, child_b bits), params, (center, 0).
template <class S> struct Sdf {
    using V = V3<S>;
    static S mn(S a, S b) { return std::fmin(a, b); }
    static S mx(S a, S b) { return std::fmax(a, b); }
    static S clamp01(S x) { return mn(mx(x, S(0)), S(1)); }
    static S sgn(S x) { return x > S(0) ? S(1) : (x < S(0) ? S(-1) : S(0)); }
    static S len2(S x, S y) { return osqrt(x * x + y * y); }
    static V vabs(V p) { return V{oabs(p.x), oabs(p.y), oabs(p.z)}; }

    // sdSphere .. sdPlane (SdfPrimitives.slang:163-195), evalLeafDist (:308-322)
    static S shape_dist(V p, uint32_t shape, const float* q) {
        switch (shape) {
            case 0: return length(p) - S(q[0]);
            case 1:
            case 2: {
                V a = vabs(p) - V{S(q[0]), S(q[1]), S(q[2])};
                S d = length(V{mx(a.x, S(0)), mx(a.y, S(0)), mx(a.z, S(0))}) + mn(mx(a.x, mx(a.y, a.z)), S(0));
                return shape == 1 ? d : d - S(q[3]);
            }
            case 3: return len2(len2(p.x, p.z) - S(q[0]), p.y) - S(q[1]);
            case 4: {
                S y = mn(mx(p.y, -S(q[0])), S(q[0]));
                return length(V{p.x, p.y - y, p.z}) - S(q[1]);
            }
            case 5: return dot(p, V{S(q[0]), S(q[1]), S(q[2])}) + S(q[3]);
            default: return S(1e30f);
        }
    }
    // sdSphereNormal .. sdPlaneNormal (:203-261), evalLeafNormal (:324-340)
    static V shape_normal(V p, uint32_t shape, const float* q) {
        const V up{S(0), S(1), S(0)};
        switch (shape) {
            case 0: {
                S l = length(p);
                return l > S(1e-12f) ? p / l : up;
            }
            case 1:
            case 2: {
                V a = vabs(p) - V{S(q[0]), S(q[1]), S(q[2])};
                V sg{sgn(p.x), sgn(p.y), sgn(p.z)};
                if (mx(a.x, mx(a.y, a.z)) > S(0)) {
                    V m{mx(a.x, S(0)), mx(a.y, S(0)), mx(a.z, S(0))};
                    S lm = length(m);
                    if (lm < S(1e-12f)) return sg * up;
                    return (m / lm) * sg;
                }
                V n{S(0), S(0), S(0)};
                if (a.x > a.y && a.x > a.z) n.x = sg.x;
                else if (a.y > a.z) n.y = sg.y;
                else n.z = sg.z;
                return n;
            }
            case 3: {
                S lxz = len2(p.x, p.z);
                if (lxz < S(1e-12f)) return V{S(0), sgn(p.y), S(0)};
                S k = S(q[0]) / lxz;
                V d = p - V{p.x * k, S(0), p.z * k};
                S ld = length(d);
                return ld > S(1e-12f) ? d / ld : up;
            }
            case 4: {
                S y = mn(mx(p.y, -S(q[0])), S(q[0]));
                V d = p - V{S(0), y, S(0)};
                S l = length(d);
                return l > S(1e-12f) ? d / l : up;
            }
            case 5: return V{S(q[0]), S(q[1]), S(q[2])};
            default: return up;
        }
    }
    // the blend factor h of opSmoothUnion / Subtract / Intersect (:276-292)
    static S blend(uint32_t op, S da, S db, S k) {
        if (op == 1) return clamp01(S(0.5f) + (S(0.5f) * (db - da)) / k);
        if (op == 2) return clamp01(S(0.5f) - (S(0.5f) * (db + da)) / k);
        return clamp01(S(0.5f) - (S(0.5f) * (db - da)) / k);
    }
    // lerp(a, b, h) = a + (b - a) h
    static S combine(uint32_t op, S da, S db, S k) {
        if (op == 1) {
            if (k <= S(0)) return mn(da, db);
            S h = blend(op, da, db, k);
            return (db + (da - db) * h) - (k * h) * (S(1) - h);
        }
        if (op == 2) {
            if (k <= S(0)) return mx(da, -db);
            S h = blend(op, da, db, k);
            return (da + (-db - da) * h) + (k * h) * (S(1) - h);
        }
        if (k <= S(0)) return mx(da, db);
        S h = blend(op, da, db, k);
        return (db + (da - db) * h) + (k * h) * (S(1) - h);
    }
    static uint32_t count(const float* c) { return std::min(f2u(c[3]), 8u); }
    // sdfClusterDist (:377-400)
    static S dist(const float* c, V p) {
        S dv[8];
        const uint32_t n = count(c);
        for (uint32_t i = 0; i < n; ++i) {
            const float* nd = c + 16 + 12 * i;
            const uint32_t op = f2u(nd[0]);
            if (op == 0) dv[i] = shape_dist(p - V{S(nd[8]), S(nd[9]), S(nd[10])}, f2u(nd[1]), nd + 4);
            else if (op >= 1 && op <= 3) dv[i] = combine(op, dv[f2u(nd[2])], dv[f2u(nd[3])], S(nd[4]));
            else if (op == 4) dv[i] = dv[f2u(nd[2])] + S(nd[4]);
            else dv[i] = S(1e30f);
        }
        return n ? dv[n - 1] : S(1e30f);
    }
    // sdfClusterNormal (:407-491)
    static V grad(const float* c, V p) {
        S dv[8];
        V gv[8];
        const V up{S(0), S(1), S(0)};
        const uint32_t n = count(c);
        for (uint32_t i = 0; i < n; ++i) {
            const float* nd = c + 16 + 12 * i;
            const uint32_t op = f2u(nd[0]);
            const S k = S(nd[4]);
            if (op == 0) {
                V pl = p - V{S(nd[8]), S(nd[9]), S(nd[10])};
                dv[i] = shape_dist(pl, f2u(nd[1]), nd + 4);
                gv[i] = shape_normal(pl, f2u(nd[1]), nd + 4);
            } else if (op >= 1 && op <= 3) {
                const S da = dv[f2u(nd[2])], db = dv[f2u(nd[3])];
                const V ga = gv[f2u(nd[2])], gb = gv[f2u(nd[3])];
                dv[i] = combine(op, da, db, k);
                if (k <= S(0)) {
                    if (op == 1) gv[i] = (da < db) ? ga : gb;
                    else if (op == 2) gv[i] = (da > -db) ? ga : -gb;
                    else gv[i] = (da > db) ? ga : gb;
                } else {
                    const S h = blend(op, da, db, k);
                    gv[i] = op == 2 ? normalize(lerp3(ga, -gb, h)) : normalize(lerp3(gb, ga, h));
                }
            } else if (op == 4) {
                dv[i] = dv[f2u(nd[2])] + k;
                gv[i] = gv[f2u(nd[2])];
            } else {
                dv[i] = S(1e30f);
                gv[i] = up;
            }
        }
        return n ? gv[n - 1] : up;
    }
    // sdfClusterTrace (:1330-1444): the slab interval of the cluster box
    // (sdfClusterAabbT, :1306-1316) clipped to [t_min, t_max], then at most
    // min(iters, 256) steps of max(|d|, eps) until |d| < eps.
    static bool march(const float* c, V ro, V rd, V inv, S t_min, S t_max, uint32_t iters, S eps, S& t_hit) {
        const S t0x = (S(c[0]) - ro.x) * inv.x, t1x = (S(c[4]) - ro.x) * inv.x;
        const S t0y = (S(c[1]) - ro.y) * inv.y, t1y = (S(c[5]) - ro.y) * inv.y;
        const S t0z = (S(c[2]) - ro.z) * inv.z, t1z = (S(c[6]) - ro.z) * inv.z;
        const S t_enter = mx(mx(mn(t0x, t1x), mn(t0y, t1y)), mn(t0z, t1z));
        const S t_exit = mn(mn(mx(t0x, t1x), mx(t0y, t1y)), mx(t0z, t1z));
        S t = mx(t_enter, t_min);
        const S t_end = mn(t_exit, t_max);
        if (t >= t_end) return false;
        const uint32_t n = std::min(iters, 256u);
        for (uint32_t i = 0; i < n; ++i) {
            const V p = ro + rd * t;
            const S ad = oabs(dist(c, p));
            if (ad < eps) {
                t_hit = t;
                return true;
            }
            t += mx(ad, eps);
            if (t >= t_end) return false;
        }
        return false;
    }
};

template <class R> inline double vd(R x) { return value_of(x); }

// traceSdfClusters (PathTrace.slang:2212-2306): every cluster in order, each
// traced up to the closest distance so far, replacing the hit only when
// strictly closer.  Float runs trace in float; double and dual runs in double
// (the clusters carry no differentiated parameter).
template <class R>
void trace_sdf(const orc_scene& S, V3<R> ro, V3<R> rd, Hit<R>& h, Ctr& c, uint32_t iters, float eps) {
    using SR = std::conditional_t<std::is_same_v<R, float>, float, double>;
    using SV = V3<SR>;
    const SV o{SR(vd(ro.x)), SR(vd(ro.y)), SR(vd(ro.z))}, d{SR(vd(rd.x)), SR(vd(rd.y)), SR(vd(rd.z))};
    auto f = [](SR x) { return oabs(x) > SR(1e-30f) ? x : SR(1e-30f); };
    const SV inv{SR(1) / f(d.x), SR(1) / f(d.y), SR(1) / f(d.z)};
    for (uint32_t i = 0; i < S.sdf_count; ++i) {
        const float* cl = &S.sdf[(size_t)i * 112];
        c.tests++;
        SR t;
        if (!Sdf<SR>::march(cl, o, d, inv, SR(1e-3f), SR(vd(h.t)), iters, SR(eps), t)) continue;
        if (!(R(t) < h.t)) continue;
        const SV n = normalize(Sdf<SR>::grad(cl, o + d * t));
        h.hit = true;
        h.t = R(t);
        h.n = V3<R>{R(n.x), R(n.y), R(n.z)};
        h.mat = f2u(cl[7]);
        h.albedo = ld3<R>(cl + 8);
        h.rough = R(cl[11]);
        h.ior = R(cl[12]);
        h.emission = V3<R>{R(0), R(0), R(0)};
        h.user_id = 0u;
        h.kind = 3u;
        h.prim = i;
    }
}

// traceScene (PathTrace.slang:2331-2693) at shutter time `time`; the mesh is
// intersected from ro_mesh = ro - its translation at that time (:2483-2487);
// SDF clusters with sdf_iters steps and surface distance sdf_eps.
template <class R>
Hit<R> trace_scene(const orc_scene& S, V3<R> ro, V3<R> rd, Ctr& c, bool brute, R time, V3<R> ro_mesh,
                   uint32_t sdf_iters = 128u, float sdf_eps = 1e-4f) {
    Hit<R> h;
    for (uint32_t i : S.planes) test_prim(S, i, ro, rd, h, c, time);
    V3<R> inv = safe_inv(rd);
    if (!S.use_abvh || brute) {
        for (uint32_t i : S.finite) test_prim(S, i, ro, rd, h, c, time);
    } else {
        uint32_t stack[64];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const BNode& nd = S.abvh.nodes[stack[--sp]];
            c.nodes++;
            if (!box_hit(ro, inv, nd.mn, nd.mx, h.t)) continue;
            if (nd.count > 0) {
                for (uint32_t k = 0; k < nd.count; ++k)
                    test_prim(S, S.abvh.perm[nd.left_first + k], ro, rd, h, c, time);
            } else {
                stack[sp++] = nd.left_first;
                stack[sp++] = nd.left_first + 1;
            }
        }
    }
    if (S.sdf_count > 0) trace_sdf(S, ro, rd, h, c, sdf_iters, sdf_eps);
    if (S.tri_count > 0) {
        R t_mesh = h.t;
        uint32_t best = 0;
        bool found = false;
        if (brute) {
            for (uint32_t tri = 0; tri < S.tri_count; ++tri) {
                c.tests++;
                TriT<R> r = tri_test(S, tri, ro_mesh, rd);
                if (r.ok && (r.t < t_mesh || (found && r.t == t_mesh && tri < best))) {
                    found = true; t_mesh = r.t; best = tri;
                }
            }
        } else {
            uint32_t stack[128];
            int sp = 0;
            stack[sp++] = 0;
            while (sp > 0) {
                const BNode& nd = S.tbvh.nodes[stack[--sp]];
                c.nodes++;
                if (!box_hit(ro_mesh, inv, nd.mn, nd.mx, t_mesh)) continue;
                if (nd.count > 0) {
                    for (uint32_t k = 0; k < nd.count; ++k) {
                        uint32_t tri = S.tbvh.perm[nd.left_first + k];
                        c.tests++;
                        TriT<R> r = tri_test(S, tri, ro_mesh, rd);
                        if (r.ok && (r.t < t_mesh || (found && r.t == t_mesh && tri < best))) {
                            found = true; t_mesh = r.t; best = tri;
                        }
                    }
                } else {
                    stack[sp++] = nd.left_first;
                    stack[sp++] = nd.left_first + 1;
                }
            }
        }
        if (found) mesh_fill(S, best, t_mesh, h);
    }
    return h;
}
template <class R> Hit<R> trace_scene(const orc_scene& S, V3<R> ro, V3<R> rd, Ctr& c, bool brute = false) {
    return trace_scene(S, ro, rd, c, brute, R(1), ro);
}

template <class R> struct PT {
    using M = typename MathOf<R>::M;
    const orc_scene& S;
    const orc_frame& F;
    Ctr& c;

    static constexpr R PI = R(3.14159265358979);

    V3<R> clamp_indirect(V3<R> v) const {
        R lim = R(F.firefly_clamp);
        if (lim <= R(0)) return v;
        R m = maxc(v);
        return (m > lim) ? v * (lim / m) : v;
    }

    V3<R> cos_sample(V3<R> n, uint32_t& seed) const {
        R u1 = randf<R>(seed);
        R u2 = randf<R>(seed);
        R r = osqrt(u1);
        R phi = R(6.28318530718) * u2;
        R x = r * M::cos(phi);
        R y = r * M::sin(phi);
        R z = osqrt(std::max(R(0), R(1) - u1));
        V3<R> ax = oabs(n.x) > R(0.9) ? V3<R>{R(0), R(1), R(0)} : V3<R>{R(1), R(0), R(0)};
        V3<R> t = normalize(cross(n, ax));
        V3<R> b = cross(n, t);
        return normalize((t * x + b * y) + n * z);
    }

    V3<R> env_texel(uint32_t x, uint32_t y) const {
        const float* p = &S.env[((size_t)y * S.W + x) * 3];
        return V3<R>{R(p[0]), R(p[1]), R(p[2])};
    }

    // ---- sky lights: the frame's sun and moon (gradient sky) ----------------
    static R sat(R x) { return clampv(x, R(0), R(1)); }
    V3<R> sun_dir() const { return V3<R>{R(F.sun[0]), R(F.sun[1]), R(F.sun[2])}; }
    V3<R> moon_dir() const { return V3<R>{R(F.moon[0]), R(F.moon[1]), R(F.moon[2])}; }

    // sampleSunDisc (PathTraceMath.slang:55-67).
    V3<R> sample_sun_disc(V3<R> dir, uint32_t& seed) const {
        R u1 = randf<R>(seed);
        R u2 = randf<R>(seed);
        R r = osqrt(u1) * R(0.00465);
        R phi = R(6.28318530718) * u2;
        R dx = M::cos(phi) * r, dy = M::sin(phi) * r;
        V3<R> ax = oabs(dir.y) < R(0.99) ? V3<R>{R(0), R(1), R(0)} : V3<R>{R(1), R(0), R(0)};
        V3<R> t = normalize(cross(ax, dir));
        V3<R> b = cross(dir, t);
        return normalize((dir + t * dx) + b * dy);
    }
    // sun_rad of the sun NEE (PathTrace.slang:6577-6600).
    static V3<R> sun_radiance(R elev) {
        R am = R(1) / std::max(elev, R(0.04));
        R trans_atm = M::exp(R(-0.30) * am);
        R warmth = sat((am - R(1)) * R(0.18));
        V3<R> col = lerp3(V3<R>{R(1.00), R(0.97), R(0.92)}, V3<R>{R(1.55), R(0.55), R(0.20)}, warmth);
        return col * (R(80) * trans_atm);
    }
    // moon_rad of the moon NEE (PathTrace.slang:6636-6654).
    static V3<R> moon_radiance(R elev, R phase_b) {
        R m_am = R(1) / std::max(elev, R(0.04));
        R m_atm = M::exp(R(-0.30) * m_am);
        R m_warmth = sat((m_am - R(1)) * R(0.18));
        V3<R> m_col = lerp3(V3<R>{R(0.92), R(0.94), R(1.00)}, V3<R>{R(1.20), R(0.65), R(0.40)}, m_warmth);
        return m_col * ((R(0.3) * phase_b) * m_atm);
    }
    // sunDisc (PathTrace.slang:3112-3192).
    V3<R> sun_disc(V3<R> rd) const {
        const V3<R> zero{R(0), R(0), R(0)};
        V3<R> sd = sun_dir();
        if (sd.y < R(-0.10)) return zero;
        R ang = dot(rd, sd);
        R sun_mult = std::max(R(F.sun_extra[0]) * R(F.sun_extra[1]), R(0.1));
        R r_deg = R(0.55) * sun_mult;
        R cos_r = M::cos(r_deg * R(0.0174533));
        R hz = sat(R(1) - sd.y);
        R halo_deg = R(2) + hz * (R(8) - R(2));
        R cos_halo = M::cos(halo_deg * R(0.0174533));
        if (ang <= cos_halo) return zero;
        R vflat = R(F.sun_extra[2]);
        bool inside = false;
        if (ang > cos_r) {
            if (!(vflat < R(0.999))) {
                inside = true;
            } else {
                V3<R> lr = normalize(cross(V3<R>{R(0), R(1), R(0)}, sd));
                V3<R> lu = cross(sd, lr);
                V3<R> delta = rd - sd * ang;
                R h = dot(delta, lr), v = dot(delta, lu);
                R rh = r_deg * R(0.0174533);
                R rv = rh * vflat;
                inside = (h * h) / (rh * rh) + (v * v) / (rv * rv) <= R(1);
            }
        }
        R am = R(1) / std::max(sd.y, R(0.04));
        R atm = M::exp(R(-0.30) * am);
        R warmth = sat((am - R(1)) * R(0.18));
        V3<R> col = lerp3(V3<R>{R(1.00), R(0.98), R(0.94)}, V3<R>{R(1.55), R(0.55), R(0.20)}, warmth);
        R t_h = (ang - cos_halo) / (R(1) - cos_halo);
        R falloff = R(8) + hz * (R(3) - R(8));
        R halo = M::pow(t_h, falloff);
        R bright = std::max(R(F.sun_disc_brightness), R(0));
        V3<R> halo_col = ((col * halo) * atm) * R(4);
        if (inside) return (col * (R(250) * atm) + halo_col) * bright;
        return halo_col * bright;
    }
    V3<R> moon_texel(int x, int y) const {
        if (S.moon.empty()) return V3<R>{R(0), R(0), R(0)};
        const float* p = &S.moon[((size_t)y * S.moon_w + x) * 4];
        return V3<R>{R(p[0]), R(p[1]), R(p[2])};
    }
    // moonDisc (PathTrace.slang:3212-3353).
    V3<R> moon_disc(V3<R> rd) const {
        const V3<R> zero{R(0), R(0), R(0)};
        V3<R> md = moon_dir();
        if (md.y < R(-0.10)) return zero;
        R ang = dot(rd, md);
        R size_mult = std::max(R(F.moon_extra[0]) * R(F.moon_extra[1]), R(0.1));
        R r_deg = R(0.55) * size_mult;
        R halo_deg = R(1.5) * size_mult;
        R cos_r = M::cos(r_deg * R(0.0174533));
        R cos_halo = M::cos(halo_deg * R(0.0174533));
        if (ang <= cos_halo) return zero;
        V3<R> tint{R(0.92), R(0.94), R(1.00)};
        R m_am = R(1) / std::max(md.y, R(0.04));
        R m_atm = M::exp(R(-0.30) * m_am);
        R warmth = sat((m_am - R(1)) * R(0.18));
        V3<R> col_atm = lerp3(V3<R>{R(1), R(1), R(1)}, V3<R>{R(1.20), R(0.65), R(0.40)}, warmth);
        R t_h = (ang - cos_halo) / (R(1) - cos_halo);
        R halo = M::pow(t_h, R(4)) * R(0.6);
        V3<R> tc = tint * col_atm;
        if (ang > cos_r) {
            R t_d = (ang - cos_r) / (R(1) - cos_r);
            V3<R> ax = oabs(md.y) < R(0.99) ? V3<R>{R(0), R(1), R(0)} : V3<R>{R(1), R(0), R(0)};
            V3<R> ta = normalize(cross(ax, md));
            V3<R> tb = cross(md, ta);
            R disc_scale = R(1) / std::max(M::sin(r_deg * R(0.0174533)), R(1e-5));
            R du = dot(rd, ta) * disc_scale;
            R dv = dot(rd, tb) * disc_scale;
            R r2 = du * du + dv * dv;
            R zz = osqrt(std::max(R(1) - r2, R(0)));
            V3<R> n_world = (ta * du + tb * dv) - md * zz;
            R lit = sat(dot(n_world, sun_dir()) * R(10) + R(0.05));
            if (lit <= R(0)) return ((tc * halo) * m_atm) * R(0.05);
            R anti_sun_ang = M::acos(sat(-dot(md, sun_dir())));
            R et = sat((anti_sun_ang - R(0.0122)) / (R(0.0218) - R(0.0122)));
            R eclipse = R(1) - (et * et) * (R(3) - R(2) * et);
            R lon = M::atan2(du, zz);
            R lat = M::asin(clampv(dv, R(-1), R(1)));
            R u_tex = sat(R(0.5) + lon / R(6.28318530718));
            R v_tex = sat(R(0.5) - lat / R(3.14159265358979));
            V3<R> surface{R(0), R(0), R(0)};
            if (!S.moon.empty()) {
                const int W = (int)S.moon_w, H = (int)S.moon_h;
                R fx = u_tex * R(W) - R(0.5), fy = v_tex * R(H) - R(0.5);
                int ix = (int)ofloor(fx), iy = (int)ofloor(fy);
                R ax2 = fx - R(ix), ay2 = fy - R(iy);
                int x0 = std::min(std::max(ix, 0), W - 1), x1 = std::min(std::max(ix + 1, 0), W - 1);
                int y0 = std::min(std::max(iy, 0), H - 1), y1 = std::min(std::max(iy + 1, 0), H - 1);
                surface = lerp3(lerp3(moon_texel(x0, y0), moon_texel(x1, y0), ax2),
                                lerp3(moon_texel(x0, y1), moon_texel(x1, y1), ax2), ay2);
            }
            R limb = R(0.55) + R(0.45) * osqrt(sat(t_d));
            R bright_mult = std::max(R(F.moon_extra[2]), R(0));
            R disc_b = R(3.5) * bright_mult;
            V3<R> lit_col = (tc * surface) * (((disc_b * lit) * limb) * m_atm) +
                            (((tc * halo) * lit) * m_atm) * (R(0.6) * bright_mult);
            V3<R> eclipse_col = (V3<R>{R(0.55), R(0.15), R(0.04)} * surface) * ((limb * m_atm) * R(1.2));
            return lerp3(lit_col, eclipse_col, eclipse);
        }
        R phase_b = R(0.5) * (R(1) - M::cos(R(F.moon[3])));
        return (((tc * halo) * phase_b) * m_atm) * R(0.3);
    }

    // isExtractedClusterDir (PathTrace.slang:3679-3696).
    bool is_cluster_dir(V3<R> d) const {
        if (!S.env_present || F.hdri_light_count == 0u) return false;
        R th = M::acos(clampv(d.y, R(-1), R(1)));
        R st = M::sin(th);
        if (st < R(1e-6)) return false;
        R ph = M::atan2(d.z, d.x);
        uint32_t W = S.W, H = S.H;
        R u = (ph + PI) * (R(1) / R(6.28318530717958));
        R v = th * (R(1) / PI);
        uint32_t ui = std::min((uint32_t)(u * R(W)), W - 1u);
        uint32_t vi = std::min((uint32_t)(v * R(H)), H - 1u);
        const float* row = S.cond.data() + (size_t)vi * W;
        float c_prev = ui > 0u ? row[ui - 1u] : 0.0f;
        return row[ui] <= c_prev;
    }

    // skyColor (PathTrace.slang:3373-3431).  Without an env map: the base
    // gradient plus the moon and sun discs (zero below y = -0.1,
    // PathTrace.slang:3113, 3213, and for no sun or moon).
    V3<R> sky(V3<R> rd) const {
        if (!S.env_present) {
            R t = clampv(R(0.5) * (rd.y + R(1)), R(0), R(1));
            V3<R> base = lerp3(V3<R>{R(0.5), R(0.55), R(0.62)}, V3<R>{R(1), R(1), R(1)}, t);
            return (base + moon_disc(rd)) + sun_disc(rd);
        }
        int W = (int)S.W, H = (int)S.H;
        R u = M::atan2(rd.z, rd.x) * (R(1) / R(6.28318530718)) + R(0.5);
        R v = M::acos(clampv(rd.y, R(-1), R(1))) * (R(1) / R(3.14159265359));
        R fx = u * R(W) - R(0.5), fy = v * R(H) - R(0.5);
        R flx = ofloor(fx), fly = ofloor(fy);
        int ix = (int)flx, iy = (int)fly;
        R ax = fx - R(ix), ay = fy - R(iy);
        int x0 = (ix % W + W) % W;
        int x1 = (x0 + 1) % W;
        int y0 = std::min(std::max(iy, 0), H - 1);
        int y1 = std::min(std::max(iy + 1, 0), H - 1);
        V3<R> c00 = env_texel(x0, y0), c10 = env_texel(x1, y0);
        V3<R> c01 = env_texel(x0, y1), c11 = env_texel(x1, y1);
        V3<R> s = lerp3(lerp3(c00, c10, ax), lerp3(c01, c11, ax), ay);
        return s * R(S.env_intensity);
    }

    struct EnvS {
        V3<R> dir, color;
        R pdf;
    };
    uint32_t cdf_search(const float* cdf, uint32_t count, R u) const {
        uint32_t lo = 0, hi = count;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (R(cdf[mid]) < u) lo = mid + 1; else hi = mid;
        }
        return std::min(lo, count - 1);
    }
    // sampleEnvMap (PathTrace.slang:3638-3670).
    EnvS sample_env(R xi0, R xi1) const {
        EnvS s{{R(0), R(1), R(0)}, {R(0), R(0), R(0)}, R(0)};
        if (!S.env_present || !(S.env_total > 0.0f)) return s;
        uint32_t W = S.W, H = S.H;
        uint32_t vi = cdf_search(S.marg.data(), H, xi1);
        uint32_t ui = cdf_search(S.cond.data() + (size_t)vi * W, W, xi0);
        R u = (R(ui) + R(0.5)) / R(W);
        R v = (R(vi) + R(0.5)) / R(H);
        R theta = v * PI;
        R phi = u * R(6.28318530717958) - PI;
        R st = M::sin(theta);
        s.dir = V3<R>{st * M::cos(phi), M::cos(theta), st * M::sin(phi)};
        V3<R> raw = env_texel(ui, vi);
        s.color = raw * R(S.env_intensity);
        R lum = (R(0.2126) * raw.x + R(0.7152) * raw.y) + R(0.0722) * raw.z;
        if (st > R(1e-6)) s.pdf = (lum * R(W * H)) / (((R(2) * PI) * PI) * R(S.env_total));
        return s;
    }
    // pdfEnvMapDir (PathTrace.slang:3710-3727).
    R pdf_env(V3<R> d) const {
        if (!S.env_present || !(S.env_total > 0.0f)) return R(0);
        R ct = clampv(d.y, R(-1), R(1));
        R th = M::acos(ct);
        R st = M::sin(th);
        if (st < R(1e-6)) return R(0);
        R ph = M::atan2(d.z, d.x);
        uint32_t W = S.W, H = S.H;
        R u = (ph + PI) * (R(1) / R(6.28318530717958));
        R v = th * (R(1) / PI);
        uint32_t ui = std::min((uint32_t)(u * R(W)), W - 1u);
        uint32_t vi = std::min((uint32_t)(v * R(H)), H - 1u);
        V3<R> raw = env_texel(ui, vi);
        R lum = (R(0.2126) * raw.x + R(0.7152) * raw.y) + R(0.0722) * raw.z;
        return (lum * R(W * H)) / (((R(2) * PI) * PI) * R(S.env_total));
    }
    static R mis(R a, R b) { return a / std::max(a + b, R(1e-12)); }

    // lightTreeImportance (PathTrace.slang:3929-3966).
    R importance(const LightTreeNode& n, V3<R> p, V3<R> nf) const {
        V3<R> mn = ld3<R>(n.mn), mx = ld3<R>(n.mx);
        V3<R> ce = (mn + mx) * R(0.5);
        V3<R> to = ce - p;
        R d2 = std::max(dot(to, to), R(1e-12));
        R d = osqrt(d2);
        V3<R> dir = to / d;
        V3<R> ext = (mx - mn) * R(0.5);
        R rad2 = std::max(dot(ext, ext), R(1e-12));
        R fall = R(1) / std::max(d2, rad2);
        R orient = R(1);
        if (R(n.cos_half) > R(-0.999)) {
            R ce2 = -dot(ld3<R>(n.axis), dir);
            R t = (ce2 - (R(n.cos_half) - R(0.2))) / R(0.2);
            orient = clampv(t, R(0), R(1));
        }
        R recv = std::max(R(0), dot(nf, dir));
        return std::max(((R(n.intensity) * fall) * orient) * recv, R(1e-8));
    }

    struct LS {
        V3<R> dir, Le;
        R pdf, dist;
    };
    // sampleAnalyticLight (PathTrace.slang:4035-4179).
    LS sample_light(V3<R> p, V3<R> nf, uint32_t& seed) const {
        LS s{{R(0), R(1), R(0)}, {R(0), R(0), R(0)}, R(0), R(1e30)};
        uint32_t n = S.light_count;
        if (n == 0) return s;
        uint32_t pick;
        R pick_pdf;
        if (F.light_tree_enabled != 0u && !S.ltree.empty()) {
            uint32_t idx = 0, lid = 0;
            pick_pdf = R(1);
            for (uint32_t d = 0; d < 32u; ++d) {
                const LightTreeNode& nd = S.ltree[idx];
                if (nd.count > 0u) { lid = nd.left_first; break; }
                uint32_t l = nd.left_first, r = l + 1;
                R il = importance(S.ltree[l], p, nf);
                R ir = importance(S.ltree[r], p, nf);
                R tot = il + ir;
                R pl = (tot > R(1e-30)) ? (il / tot) : R(0.5);
                R u = randf<R>(seed);
                if (u < pl) { idx = l; pick_pdf *= pl; } else { idx = r; pick_pdf *= (R(1) - pl); }
            }
            pick = std::min(lid, n - 1);
        } else {
            R u = randf<R>(seed);
            pick = std::min((uint32_t)(u * R(n)), n - 1);
            pick_pdf = R(1) / R(n);
        }
        const float* L = &S.lights[(size_t)pick * 20];
        uint32_t type = (uint32_t)(L[7] + 0.5f);
        V3<R> pos = ld3<R>(L);
        R radius = R(L[3]);
        V3<R> I = ld3<R>(L + 4);
        R q[4] = {R(L[16]), R(L[17]), R(L[18]), R(L[19])};
        if (type == LIGHT_POINT) {
            V3<R> tl = pos - p;
            R d2 = std::max(dot(tl, tl), R(1e-12));
            R d = osqrt(d2);
            s.dir = tl / d; s.Le = I / d2; s.pdf = pick_pdf; s.dist = d;
            return s;
        }
        if (type == LIGHT_SPOT) {
            V3<R> tl = pos - p;
            R d2 = std::max(dot(tl, tl), R(1e-12));
            R d = osqrt(d2);
            V3<R> tld = tl / d;
            V3<R> axis = normalize(quat_rotate(q, ld3<R>(L + 8)));
            R cd = dot(-tld, axis);
            R ci = R(L[12]), co = R(L[11]);
            R fo = R(0);
            if (cd >= ci) fo = R(1);
            else if (cd > co) fo = (cd - co) / std::max(ci - co, R(1e-6));
            s.dir = tld; s.Le = I * (fo / d2); s.pdf = pick_pdf; s.dist = d;
            return s;
        }
        if (type == LIGHT_SPHERE) {
            R x0 = randf<R>(seed);
            R x1 = randf<R>(seed);
            R z = R(1) - R(2) * x0;
            R r = osqrt(std::max(R(0), R(1) - z * z));
            R phi = R(6.28318530717958) * x1;
            V3<R> nl{r * M::cos(phi), r * M::sin(phi), z};
            V3<R> lp = pos + nl * radius;
            V3<R> tl = lp - p;
            R d2 = std::max(dot(tl, tl), R(1e-12));
            R d = osqrt(d2);
            V3<R> wi = tl / d;
            R cl = std::max(R(0), dot(nl, -wi));
            if (cl <= R(0)) return s;
            R area = ((R(4) * PI) * radius) * radius;
            R pdf_sa = d2 / std::max(area * cl, R(1e-12));
            s.dir = wi; s.Le = I; s.pdf = pick_pdf * pdf_sa; s.dist = d;
            return s;
        }
        if (type == 4u) {   // emissive triangle, uniform in area, one-sided
            R x0 = randf<R>(seed);
            R x1 = randf<R>(seed);
            R su = osqrt(x0);
            R b0 = R(1) - su;
            R b1 = x1 * su;
            V3<R> a = pos, b = ld3<R>(L + 8), cc = ld3<R>(L + 12);
            V3<R> lp = (a * b0 + b * b1) + cc * ((R(1) - b0) - b1);
            V3<R> cr = cross(b - a, cc - a);
            R area = R(0.5) * length(cr);
            V3<R> nt = normalize(cr);
            V3<R> tl = lp - p;
            R d2 = std::max(dot(tl, tl), R(1e-12));
            R d = osqrt(d2);
            V3<R> wi = tl / d;
            R cl = dot(nt, -wi);
            if (cl <= R(0)) return s;
            R pdf_sa = d2 / std::max(area * cl, R(1e-12));
            s.dir = wi; s.Le = I; s.pdf = pick_pdf * pdf_sa; s.dist = d;
            return s;
        }
        if (type == LIGHT_QUAD) {
            R x0 = randf<R>(seed) * R(2) - R(1);
            R x1 = randf<R>(seed) * R(2) - R(1);
            V3<R> uv = quat_rotate(q, ld3<R>(L + 12));
            R ul = length(uv);
            if (ul < R(1e-6)) return s;
            V3<R> nq = normalize(quat_rotate(q, ld3<R>(L + 8)));
            V3<R> va = normalize(cross(nq, uv));
            V3<R> lp = (pos + uv * x0) + va * (R(L[15]) * x1);
            V3<R> tl = lp - p;
            R d2 = std::max(dot(tl, tl), R(1e-12));
            R d = osqrt(d2);
            V3<R> wi = tl / d;
            R cl = dot(nq, -wi);
            if (cl <= R(0)) return s;
            R area = (R(4) * ul) * R(L[15]);
            R pdf_sa = d2 / std::max(area * cl, R(1e-12));
            s.dir = wi; s.Le = I; s.pdf = pick_pdf * pdf_sa; s.dist = d;
            return s;
        }
        return s;
    }

    // The closest hit at shutter time `time`, the mesh translated by
    // lerp(mesh_shift_prev, mesh_shift_curr, time) when its translation is on
    // (PathTrace.slang:2483-2487).
    Hit<R> trace(V3<R> ro, V3<R> rd, R time) const {
        V3<R> rm = ro;
        if (F.mesh_shift_active) rm = ro - mix3(ld3<R>(F.mesh_shift_prev), ld3<R>(F.mesh_shift_curr), time);
        return trace_scene(S, ro, rd, c, false, time, rm, F.sdf_max_iters, F.sdf_epsilon);
    }

    // transmittance (PathTrace.slang:5089-5196), no clouds; water absorbs
    // along the segments the shadow ray runs in it (start_in_water: the
    // query starts under water) and refracts it like glass, F0 0.02.
    V3<R> transmittance(V3<R> o, V3<R> d, R t_max, R time, bool start_in_water = false) const {
        int kmax = (int)F.refract_bounces;
        V3<R> t{R(1), R(1), R(1)};
        bool sr_in_water = start_in_water;
        for (int i = 0; i <= kmax; ++i) {
            c.shadow++;
            Hit<R> h = trace(o, d, time);
            if (sr_in_water) t = t * absorb_over(std::max(h.hit ? std::min(h.t, t_max) : t_max, R(0)));
            if (!h.hit || h.t >= t_max) return t;
            if (h.mat != MAT_DIELECTRIC && h.mat != MAT_WATER) return V3<R>{R(0), R(0), R(0)};
            if (i == kmax) return V3<R>{R(0), R(0), R(0)};
            R ndi = dot(h.n, d);
            V3<R> nf = (ndi < R(0)) ? h.n : -h.n;
            const R ior = (h.mat == MAT_WATER) ? water_ior_of(h.ior) : h.ior;
            R ia = (ndi < R(0)) ? R(1) : ior;
            R ib = (ndi < R(0)) ? ior : R(1);
            R eta = ia / ib;
            R ci = oabs(dot(nf, -d));
            R q = (ia - ib) / (ia + ib);
            R r0 = (h.mat == MAT_WATER) ? R(0.02) : q * q;
            R m = R(1) - ci;
            R fr = r0 + ((((R(1) - r0) * m) * m * m) * m) * m;
            V3<R> rf = refract(d, nf, eta);
            if (length(rf) < R(1e-3)) return V3<R>{R(0), R(0), R(0)};
            if (h.mat == MAT_WATER) {
                t = t * (R(1) - fr);
                sr_in_water = !sr_in_water;
            } else {
                t = t * (h.albedo * (R(1) - fr));
            }
            R adv = h.t + R(1e-3);
            o = o + d * adv;
            d = normalize(rf);
            t_max = std::max(t_max - adv, R(0));
        }
        return V3<R>{R(0), R(0), R(0)};
    }

    // cauchyIOR (PathTraceMath.slang:111-119): n = B + C / lambda^2 through
    // (base, 587.6 nm), lambda in micrometres.
    static R cauchy_ior(R base, R cc, R lambda_nm) {
        R lum = lambda_nm * R(0.001);
        R l2 = lum * lum;
        R ld2 = R(0.5876) * R(0.5876);
        R B = base - cc / ld2;
        return B + cc / l2;
    }
    // wymanGauss, cieXYZFit, xyzToLinearSRGB, spectralWeightSRGB
    // (PathTraceMath.slang:140-192): the Wyman-Sloan-Shirley fit of the CIE
    // 1931 matching functions, to linear sRGB, over the band means.
    static R wyman(R x, R mu, R s1, R s2) {
        R t = (x - mu) * ((x < mu) ? s1 : s2);
        return M::exp((R(-0.5) * t) * t);
    }
    static V3<R> spectral_weight(R l) {
        R X = (R(1.056) * wyman(l, R(599.8), R(0.0264), R(0.0323)) + R(0.362) * wyman(l, R(442.0), R(0.0624), R(0.0374)))
              - R(0.065) * wyman(l, R(501.1), R(0.0490), R(0.0382));
        R Y = R(0.821) * wyman(l, R(568.8), R(0.0213), R(0.0247)) + R(0.286) * wyman(l, R(530.9), R(0.0613), R(0.0322));
        R Z = R(1.217) * wyman(l, R(437.0), R(0.0845), R(0.0278)) + R(0.681) * wyman(l, R(459.0), R(0.0385), R(0.0725));
        R r = (R(3.2404542) * X - R(1.5371385) * Y) - R(0.4985314) * Z;
        R g = (R(-0.9692660) * X + R(1.8760108) * Y) + R(0.0415560) * Z;
        R b = (R(0.0556434) * X - R(0.2040259) * Y) + R(1.0572252) * Z;
        return V3<R>{r * R(2.730792), g * R(3.443936), b * R(3.606855)};
    }

    // ptConcentricDisk (PathTraceMath.slang:223-236).
    static void concentric_disk(R u1, R u2, R& x, R& y) {
        R a = R(2) * u1 - R(1);
        R b = R(2) * u2 - R(1);
        R r, phi;
        const R qpi = R(3.14159265358979) / R(4), hpi = R(3.14159265358979) / R(2);
        if (oabs(a) > oabs(b)) {
            r = a;
            phi = qpi * (b / std::max(oabs(a), R(1e-6)));
        } else {
            r = b;
            phi = hpi - qpi * (a / std::max(oabs(b), R(1e-6)));
        }
        x = r * M::cos(phi);
        y = r * M::sin(phi);
    }
    // ptPolygonSample (PathTraceMath.slang:268-286): a wedge of the regular
    // N-gon picked by u1, then the sqrt-barycentric point of that triangle.
    static void polygon_sample(R u1, R u2, int blades, R rot, R& x, R& y) {
        R n = R(blades);
        R wedge = R(6.28318530717958) / n;
        R kf = ofloor(u1 * n);
        R ku = u1 * n - kf;
        R a0 = rot + kf * wedge;
        R a1 = a0 + wedge;
        R v1x = M::cos(a0), v1y = M::sin(a0);
        R v2x = M::cos(a1), v2y = M::sin(a1);
        R su = osqrt(ku);
        R w1 = R(1) - u2;
        x = su * (w1 * v1x + u2 * v2x);
        y = su * (w1 * v1y + u2 * v2y);
    }
    // ptCatEye (PathTraceMath.slang:296-310): the radial component of the lens
    // sample narrowed with the image eccentricity |uv|.
    static void cat_eye(R& x, R& y, R ux, R uy, R strength) {
        R len = osqrt(ux * ux + uy * uy);
        R ecc = clampv(len, R(0), R(1));
        R squeeze = strength * ecc;
        if (squeeze < R(1e-5)) return;
        R rx = R(1), ry = R(0);
        if (ecc > R(1e-5)) { rx = ux / len; ry = uy / len; }
        R tx = -ry, ty = rx;
        R ar = x * rx + y * ry;
        R at = x * tx + y * ty;
        ar = ar * (R(1) - R(0.7) * squeeze);
        x = rx * ar + tx * at;
        y = ry * ar + ty * at;
    }

    // Camera ray of sample s of pixel (x, y) (PathTrace.slang:5729-5816):
    // the jitter, then, each only when its feature is on, the hero wavelength
    // (stratified over the samples), the shutter time and the lens point.
    void camera(uint32_t x, uint32_t y, uint32_t s, uint32_t spp, uint32_t& seed, V3<R>& ro, V3<R>& rd, R& time,
                R& lambda) const {
        const R w = R(F.width), hgt = R(F.height);
        R fovtan = R(F.pos_fovtan[3]), aspect = R(F.fwd_aspect[3]);
        V3<R> fwd{R(F.fwd_aspect[0]), R(F.fwd_aspect[1]), R(F.fwd_aspect[2])};
        V3<R> rt{R(F.right_xyz[0]), R(F.right_xyz[1]), R(F.right_xyz[2])};
        V3<R> up{R(F.up_xyz[0]), R(F.up_xyz[1]), R(F.up_xyz[2])};
        R jx, jy;
        if ((F.flags & 2u) && s == 0u) {   // the denoiser's fixed offset (PathTrace.slang:5730-5731)
            jx = R(F.jitter[0]);
            jy = R(F.jitter[1]);
        } else {
            jx = randf<R>(seed);
            jy = randf<R>(seed);
            jx = jx - R(0.5);
            jy = jy - R(0.5);
        }
        R ux = (((R(x) + R(0.5)) + jx) / w) * R(2) - R(1);
        R uy = (((R(y) + R(0.5)) + jy) / hgt) * R(2) - R(1);
        uy = -uy;
        lambda = R(-1);
        if (F.spectral) {
            R u = (R(s) + randf<R>(seed)) / R(spp);
            lambda = R(380) + u * (R(730) - R(380));
        }
        time = R(1);
        if (F.motion_blur) time = randf<R>(seed);
        ro = V3<R>{R(F.pos_fovtan[0]), R(F.pos_fovtan[1]), R(F.pos_fovtan[2])};
        rd = normalize((fwd + rt * ((ux * aspect) * fovtan)) + up * (uy * fovtan));
        if (F.dof[0] > 0.0f) {
            const R aperture = R(F.dof[0]), focal = R(F.dof[1]);
            const int blades = (int)(F.dof[2] + 0.5f);
            R u1 = randf<R>(seed);
            R u2 = randf<R>(seed);
            R ax, ay;
            if (blades >= 3) polygon_sample(u1, u2, blades, R(F.dof[3]), ax, ay);
            else concentric_disk(u1, u2, ax, ay);
            if (F.cat_eye > 0.0f) cat_eye(ax, ay, ux, uy, R(F.cat_eye));
            V3<R> lens = rt * (ax * aperture) + up * (ay * aperture);
            V3<R> fp = ro + rd * focal;
            ro = ro + lens;
            rd = normalize(fp - ro);
        }
    }

    // ---- Water (PathTrace.slang:5207-5270, PathTraceCloud.slang:31-69) ----
    // cloud_hash: lowbias32 of the integer lattice point, 24-bit uniform.
    static uint32_t lowbias32(uint32_t x) {
        x ^= x >> 16; x *= 0x7feb352du;
        x ^= x >> 15; x *= 0x846ca68bu;
        x ^= x >> 16;
        return x;
    }
    static R lattice(R x, R y, R z) {
        uint32_t h = lowbias32((uint32_t)(int32_t)value_of(x));
        h = lowbias32(h ^ (uint32_t)(int32_t)value_of(y));
        h = lowbias32(h ^ (uint32_t)(int32_t)value_of(z));
        return R(float(h & 0x00FFFFFFu) * (1.0f / 16777216.0f));
    }
    static R lerpf(R a, R b, R t) { return a + t * (b - a); }
    // cloud_value_noise: quintic-smoothed trilinear value noise.
    static R value_noise(V3<R> p) {
        const V3<R> i{ofloor(p.x), ofloor(p.y), ofloor(p.z)};
        V3<R> f = p - i;
        auto q = [](R t) { return ((t * t) * t) * (t * (t * R(6) - R(15)) + R(10)); };
        f = V3<R>{q(f.x), q(f.y), q(f.z)};
        const R a = lattice(i.x, i.y, i.z), b = lattice(i.x + R(1), i.y, i.z);
        const R c = lattice(i.x, i.y + R(1), i.z), d = lattice(i.x + R(1), i.y + R(1), i.z);
        const R e = lattice(i.x, i.y, i.z + R(1)), fx = lattice(i.x + R(1), i.y, i.z + R(1));
        const R gx = lattice(i.x, i.y + R(1), i.z + R(1)), hx = lattice(i.x + R(1), i.y + R(1), i.z + R(1));
        return lerpf(lerpf(lerpf(a, b, f.x), lerpf(c, d, f.x), f.y), lerpf(lerpf(e, fx, f.x), lerpf(gx, hx, f.x), f.y),
                     f.z);
    }
    // waterShadingNormal: the macro normal bent by the central-difference
    // gradient of two drifting octaves of value noise over the xz plane.
    V3<R> water_normal(V3<R> wp, V3<R> n) const {
        const R scale = R(F.wave[0]), amp = R(F.wave[1]), speed = R(F.wave[2]);
        const R t = R(F.wave[3]) * speed;
        if (amp <= R(0) || scale <= R(0)) return n;
        const R eps = R(0.15);
        const R px = wp.x * scale, pz = wp.z * scale;
        const R dax = R(0.13) * t, daz = R(0.07) * t;
        const R dbx = R(-0.09) * t, dbz = R(0.11) * t;
        const R k = R(2.02);
        auto h = [&](R x, R z, R x2, R z2) {
            return value_noise(V3<R>{x + dax, R(0), z + daz}) + R(0.5) * value_noise(V3<R>{x2 + dbx, R(0), z2 + dbz});
        };
        const R n_px = h(px + eps, pz, px * k + eps, pz * k);
        const R n_nx = h(px - eps, pz, px * k - eps, pz * k);
        const R n_pz = h(px, pz + eps, px * k, pz * k + eps);
        const R n_nz = h(px, pz - eps, px * k, pz * k - eps);
        const R dhx = (n_px - n_nx) / (R(2) * eps);
        const R dhz = (n_pz - n_nz) / (R(2) * eps);
        const V3<R> t1 = (oabs(n.y) < R(0.99)) ? V3<R>{R(0), R(1), R(0)} : V3<R>{R(1), R(0), R(0)};
        const V3<R> tx = normalize(cross(t1, n));
        const V3<R> tz = cross(n, tx);
        return normalize(n - (tx * dhx + tz * dhz) * amp);
    }
    V3<R> absorb_over(R seg) const {
        return V3<R>{M::exp(-R(F.water_absorb[0]) * seg), M::exp(-R(F.water_absorb[1]) * seg),
                     M::exp(-R(F.water_absorb[2]) * seg)};
    }
    R water_ior_of(R h_ior) const { return clampv((h_ior > R(0)) ? h_ior : R(F.water_ior), R(1), R(2.4)); }

    // ---- PBR material atlas (PathTrace.slang:1435-1605, applyPbrTextures;
    // the software device's CPU mirror SoftwareTracer.cpp:356-525) ----
    static constexpr uint32_t kPbrTile = 256u, kPbrMaxLod = 8u;
    struct V4 {
        R x, y, z, w;
    };
    static V4 mix4(V4 a, V4 b, R t) { return {a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z), a.w + t * (b.w - a.w)}; }
    V4 texel(uint32_t x, uint32_t y) const {
        const float* q = &S.atlas[((size_t)y * S.atlas_w + x) * 4];
        return {R(q[0]), R(q[1]), R(q[2]), R(q[3])};
    }
    // pbrMipRegion + samplePbrTileLevel: bilinear within mip `level` of the
    // tile, wrapped at the level's resolution.
    V4 tile_level(uint32_t tile, R u, R v, uint32_t level) const {
        const uint32_t row_base = tile * kPbrTile;
        uint32_t x0, y0, dim;
        if (level == 0u) {
            x0 = 0u; y0 = row_base; dim = kPbrTile;
        } else {
            dim = kPbrTile >> level;
            x0 = kPbrTile;
            y0 = row_base + (kPbrTile - 2u * dim);
        }
        const R fu = u - ofloor(u), fv = v - ofloor(v);
        const R fx = fu * R(dim) - R(0.5), fy = fv * R(dim) - R(0.5);
        const R fx0 = ofloor(fx), fy0 = ofloor(fy);
        const R tx = fx - fx0, ty = fy - fy0;
        const int D = (int)dim;
        int ix0 = (int)value_of(fx0), iy0 = (int)value_of(fy0);
        const int ix1 = ((ix0 + 1) % D + D) % D, iy1 = ((iy0 + 1) % D + D) % D;
        ix0 = (ix0 % D + D) % D;
        iy0 = (iy0 % D + D) % D;
        const V4 c00 = texel(x0 + (uint32_t)ix0, y0 + (uint32_t)iy0), c10 = texel(x0 + (uint32_t)ix1, y0 + (uint32_t)iy0);
        const V4 c01 = texel(x0 + (uint32_t)ix0, y0 + (uint32_t)iy1), c11 = texel(x0 + (uint32_t)ix1, y0 + (uint32_t)iy1);
        return mix4(mix4(c00, c10, tx), mix4(c01, c11, tx), ty);
    }
    // samplePbrTileLinear: trilinear between the two bracketing levels.
    V4 tile_linear(uint32_t tile, R u, R v, R lod) const {
        const R l = clampv(lod, R(0), R(kPbrMaxLod));
        const uint32_t l0 = (uint32_t)value_of(ofloor(l));
        const uint32_t l1 = std::min(l0 + 1u, kPbrMaxLod);
        const R f = l - R(l0);
        const V4 a = tile_level(tile, u, v, l0);
        if (f <= R(0)) return a;
        return mix4(a, tile_level(tile, u, v, l1), f);
    }
    static R srgb_to_linear(R c) {
        const R lo = c / R(12.92);
        const R hi = M::pow(std::max((c + R(0.055)) / R(1.055), R(0)), R(2.4));
        return (c <= R(0.04045)) ? lo : hi;
    }
    // pbrConeTexLod: the ray cone's footprint projected onto the surface.
    static R cone_lod(R cone_width, V3<R> rd, V3<R> n, R uv_scale) {
        if (uv_scale <= R(0)) return R(0);
        const R cos_i = std::max(oabs(dot(rd, n)), R(1e-3));
        const R texels = ((cone_width / cos_i) * uv_scale) * R(kPbrTile);
        return clampv(M::log2(std::max(texels, R(1))), R(0), R(kPbrMaxLod));
    }
    // applyNormalMap: Duff et al.'s branchless basis around n.
    static V3<R> normal_map(V3<R> n, V3<R> ts) {
        const R s = (n.z >= R(0)) ? R(1) : R(-1);
        const R a = R(-1) / (s + n.z);
        const R b = (n.x * n.y) * a;
        const V3<R> t{R(1) + ((s * n.x) * n.x) * a, s * b, -(s * n.x)};
        const V3<R> bt{b, s + (n.y * n.y) * a, -n.y};
        return normalize((t * ts.x + bt * ts.y) + n * ts.z);
    }
    // The hit's UV, UV density and tiles (testAnalyticPrim's sphere and plane
    // maps, PathTrace.slang:2104-2158; the mesh's barycentric UVs, :2536-2543),
    // its ray-cone LOD, then the maps applied in place: albedo (sRGB decoded,
    // times the flat tint), roughness, metallic (Lambert -> metal above 0.5),
    // the normal map last.  A tile applies when it is below the atlas's tile
    // count (the CPU mirror's gate; kPbrNoTexTile never is).
    void pbr_uv(const Hit<R>& h, V3<R> ro, V3<R> rd, R time, uint32_t tiles[4], R& u, R& v, R& uv_scale) const {
        for (int k = 0; k < 4; ++k) tiles[k] = 0xFFFFFFFFu;
        u = R(0); v = R(0); uv_scale = R(0);
        if (h.kind == 1u) {
            const float* p = &S.prims[(size_t)h.prim * 28];
            for (int k = 0; k < 4; ++k) tiles[k] = f2u(p[24 + k]);
            if ((uint32_t)p[8] == PRIM_SPHERE) {
                const V3<R> sn = h.n;
                u = R(0.5) + M::atan2(sn.z, sn.x) / (R(2) * PI);
                v = R(0.5) - M::asin(clampv(sn.y, R(-1), R(1))) / PI;
                const R sr = std::max(R(p[3]), R(1e-4));
                const R cos_lat = std::max(osqrt(sn.x * sn.x + sn.z * sn.z), R(0.05));
                uv_scale = std::max(R(1) / (PI * sr), R(1) / (((R(2) * PI) * sr) * cos_lat));
            } else {
                const V3<R> hp = ro + rd * h.t;
                const V3<R> an{oabs(h.n.x), oabs(h.n.y), oabs(h.n.z)};
                if (an.y >= an.x && an.y >= an.z) { u = hp.x; v = hp.z; }
                else if (an.x >= an.z) { u = hp.z; v = hp.y; }
                else { u = hp.x; v = hp.y; }
                uv_scale = R(1);
            }
        } else if (h.kind == 2u && !S.uvs.empty()) {
            for (int k = 0; k < 4; ++k) tiles[k] = (F.flags & 4u) ? F.mesh_tex[k] : 0xFFFFFFFFu;   // DMT_FRAME_MESH_TEXTURES
            V3<R> rm = ro;
            if (F.mesh_shift_active) rm = ro - mix3(ld3<R>(F.mesh_shift_prev), ld3<R>(F.mesh_shift_curr), time);
            R bu, bv;
            tri_bary(S, h.prim, rm, rd, bu, bv);
            const uint32_t* ix = &S.idx[(size_t)h.prim * 3];
            const float* q0 = &S.uvs[(size_t)ix[0] * 2];
            const float* q1 = &S.uvs[(size_t)ix[1] * 2];
            const float* q2 = &S.uvs[(size_t)ix[2] * 2];
            const R w0 = (R(1) - bu) - bv;
            u = (R(q0[0]) * w0 + R(q1[0]) * bu) + R(q2[0]) * bv;
            v = (R(q0[1]) * w0 + R(q1[1]) * bu) + R(q2[1]) * bv;
        }
    }
    void apply_pbr(Hit<R>& h, V3<R> ro, V3<R> rd, R time, R cone_width, R* lod_out = nullptr) const {
        uint32_t tiles[4];
        R u, v, uv_scale;
        pbr_uv(h, ro, rd, time, tiles, u, v, uv_scale);
        const uint32_t n = S.atlas_tiles;
        if (lod_out) *lod_out = cone_lod(cone_width, rd, h.n, uv_scale);
        if (!(tiles[0] < n || tiles[1] < n || tiles[2] < n || tiles[3] < n)) return;
        const R lod = cone_lod(cone_width, rd, h.n, uv_scale);
        if (tiles[0] < n) {
            const V4 a = tile_linear(tiles[0], u, v, lod);
            h.albedo = V3<R>{h.albedo.x * srgb_to_linear(a.x), h.albedo.y * srgb_to_linear(a.y),
                             h.albedo.z * srgb_to_linear(a.z)};
        }
        if (tiles[2] < n) h.rough = sat(h.rough * tile_linear(tiles[2], u, v, lod).x);
        if (tiles[3] < n) {
            const R m = tile_linear(tiles[3], u, v, lod).x;
            if (m > R(0.5) && h.mat == MAT_LAMBERT) h.mat = MAT_METAL;
        }
        if (tiles[1] < n) {
            const V4 t = tile_linear(tiles[1], u, v, lod);
            h.n = normal_map(h.n, V3<R>{t.x * R(2) - R(1), t.y * R(2) - R(1), t.z * R(2) - R(1)});
        }
    }

    // One path sample: PathTrace.slang:5803-6914 with the off-path features
    // (water, clouds, PBR, clearcoat, SSS, ReSTIR) disabled; `time` is the
    // sample's shutter time, `lambda` its hero wavelength (< 0: RGB).
    V3<R> path(V3<R> ro, V3<R> rd, uint32_t& seed, R time = R(1), R lambda = R(-1)) const {
        V3<R> T{R(1), R(1), R(1)}, L{R(0), R(0), R(0)};
        bool skip_sky = false, prev_lambert = false;
        R prev_pdf = R(0);
        const bool mis_on = F.mis_enabled != 0u;
        // ray-cone LOD state (PathTrace.slang:5913-5919)
        const R cone_spread = (R(2) * R(F.pos_fovtan[3])) / R(std::max(F.height, 1u));
        R cone_width = R(0);
        // the ray's medium (:5872): the camera's in-water flag, flipped by
        // each refraction through a water surface
        bool in_water = F.camera_in_water != 0u;
        for (uint32_t b = 0; b < F.max_bounces; ++b) {
            if (b == 0) c.primary++; else c.extension++;
            Hit<R> h = trace(ro, rd, time);
            // Beer's law over the segment just traced (:6028-6031)
            if (in_water && h.hit) T = T * absorb_over(h.t);
            if (!h.hit && in_water) {
                // escaped under water: the deep-water tint (:6046-6052)
                const V3<R> v = T * ld3<R>(F.deep_color);
                L = L + ((b == 0) ? v : clamp_indirect(v));
                break;
            }
            if (!h.hit) {
                if (mis_on && prev_lambert) {
                    // cluster directions are the clusters' own NEE's (:6074);
                    // transmittance(ro, rd, 1e30) re-traces this same missed
                    // ray and returns exactly 1: not re-traced, not counted.
                    if (!is_cluster_dir(rd)) {
                        R pe = pdf_env(rd);
                        R w = mis(prev_pdf, pe);
                        L = L + clamp_indirect((T * sky(rd)) * w);
                    }
                } else if (!skip_sky) {
                    if (b == 0) L = L + T * sky(rd);
                    else L = L + clamp_indirect(T * sky(rd));
                }
                break;
            }
            V3<R> hp = ro + rd * h.t;
            if (!S.atlas.empty()) {
                cone_width = cone_width + cone_spread * h.t;   // :6162-6171
                apply_pbr(h, ro, rd, time, cone_width);
            }
            R ndi = dot(h.n, rd);
            V3<R> nf = (ndi < R(0)) ? h.n : -h.n;
            if (any_pos(h.emission)) {
                if (b == 0) L = L + T * h.emission;
                else L = L + clamp_indirect(T * h.emission);
            }
            if (h.mat == MAT_LAMBERT) {
                if (S.light_count > 0) {
                    LS ls = sample_light(hp, nf, seed);
                    R ndl = dot(nf, ls.dir);
                    if (ls.pdf > R(0) && ndl > R(0)) {
                        R tm = std::max(ls.dist - R(1e-3), R(0));
                        V3<R> tr = transmittance(hp + nf * R(1e-3), ls.dir, tm, time, in_water);
                        if (any_pos(tr)) {
                            V3<R> brdf = h.albedo / PI;
                            V3<R> cc = ((((T * ls.Le) * tr) * brdf) * ndl) / ls.pdf;
                            L = L + ((b == 0) ? cc : clamp_indirect(cc));
                        }
                    }
                }
                if (S.env_present) {
                    R x0 = randf<R>(seed);
                    R x1 = randf<R>(seed);
                    EnvS e = sample_env(x0, x1);
                    R ndl = dot(nf, e.dir);
                    if (e.pdf > R(0) && ndl > R(0)) {
                        V3<R> tr = transmittance(hp + nf * R(1e-3), e.dir, R(1e30), time, in_water);
                        if (any_pos(tr)) {
                            V3<R> brdf = h.albedo / PI;
                            R pb = ndl * (R(1) / PI);
                            bool fire = mis_on && (b + 1u) < F.max_bounces;
                            R w = fire ? mis(e.pdf, pb) : R(1);
                            V3<R> cc = (((((T * e.color) * tr) * brdf) * ndl) * w) / e.pdf;
                            L = L + clamp_indirect(cc);
                        }
                    }
                    if (F.hdri_light_count > 0u) {
                        // sampleHdriLight (:3749-3768) and its NEE (:6503-6518)
                        R u = randf<R>(seed);
                        R acc = R(0);
                        uint32_t sel = 0u;
                        for (uint32_t i = 0; i < 8u; ++i) {
                            if (i >= F.hdri_light_count) break;
                            acc = acc + R(F.hdri_dir[i][3]);
                            sel = i;
                            if (u <= acc) break;
                        }
                        V3<R> hd = ld3<R>(F.hdri_dir[sel]);
                        V3<R> hc = ld3<R>(F.hdri_col[sel]) * R(S.env_intensity);
                        R hpdf = std::max(R(F.hdri_dir[sel][3]), R(1e-6));
                        R ndl = dot(nf, hd);
                        if (ndl > R(0) && hpdf > R(0)) {
                            V3<R> tr = transmittance(hp + nf * R(1e-3), hd, R(1e30), time, in_water);
                            if (any_pos(tr)) {
                                V3<R> brdf = h.albedo / PI;
                                V3<R> cc = ((((T * hc) * tr) * brdf) * ndl) / hpdf;
                                L = L + ((b == 0) ? cc : clamp_indirect(cc));
                            }
                        }
                    }
                    skip_sky = !mis_on;
                } else {
                    // No env map: the sun's and the moon's NEE when they are up
                    // (:6552-6660), unclamped, then the base sky as ambient.
                    V3<R> sd = sun_dir();
                    if (sd.y > R(0)) {
                        R ndl = std::max(R(0), dot(nf, sd));
                        V3<R> jit = sample_sun_disc(sd, seed);
                        if (ndl > R(0)) {
                            V3<R> tr = transmittance(hp + nf * R(1e-3), jit, R(1e30), time, in_water);
                            if (any_pos(tr)) {
                                V3<R> brdf = h.albedo / PI;
                                L = L + (((T * sun_radiance(sd.y)) * tr) * brdf) * ndl;
                            }
                        }
                    }
                    V3<R> md = moon_dir();
                    if (md.y > R(0)) {
                        R phase_b = R(0.5) * (R(1) - M::cos(R(F.moon[3])));
                        if (phase_b > R(0.01)) {
                            V3<R> jit = sample_sun_disc(md, seed);
                            R ndl = std::max(R(0), dot(nf, jit));
                            if (ndl > R(0)) {
                                V3<R> tr = transmittance(hp + nf * R(1e-3), jit, R(1e30), time, in_water);
                                if (any_pos(tr)) {
                                    V3<R> brdf = h.albedo / PI;
                                    L = L + (((T * moon_radiance(md.y, phase_b)) * tr) * brdf) * ndl;
                                }
                            }
                        }
                    }
                    R t = clampv(R(0.5) * (nf.y + R(1)), R(0), R(1));
                    V3<R> a{R(0.5), R(0.55), R(0.62)}, one{R(1), R(1), R(1)};
                    L = L + (T * h.albedo) * lerp3(a, one, t);
                    skip_sky = true;
                }
                T = T * h.albedo;
                ro = hp + nf * R(1e-3);
                rd = cos_sample(nf, seed);
                if (mis_on && S.env_present) {
                    prev_pdf = std::max(R(0), dot(nf, rd)) * (R(1) / PI);
                    prev_lambert = true;
                } else {
                    prev_lambert = false;
                }
            } else if (h.mat == MAT_METAL) {
                V3<R> r = reflect(rd, nf);
                V3<R> cs = cos_sample(nf, seed);
                r = normalize(r + (cs - nf) * h.rough);
                T = T * h.albedo;
                ro = hp + nf * R(1e-3);
                rd = r;
                skip_sky = false;
                prev_lambert = false;
            } else if (h.mat == MAT_DIELECTRIC) {
                // the hero wavelength's own index (PathTrace.slang:6721-6724)
                R ior = (lambda >= R(0)) ? cauchy_ior(h.ior, R(F.cauchy_c), lambda) : h.ior;
                R ia = (ndi < R(0)) ? R(1) : ior;
                R ib = (ndi < R(0)) ? ior : R(1);
                R eta = ia / ib;
                R ci = oabs(dot(nf, -rd));
                R r0 = (ia - ib) / (ia + ib);
                r0 = r0 * r0;
                R m = R(1) - ci;
                R fr = r0 + ((((R(1) - r0) * m) * m * m) * m) * m;
                V3<R> refl = reflect(rd, nf);
                V3<R> refr = refract(rd, nf, eta);
                bool tir = length(refr) < R(1e-3);
                bool take_refl = tir;
                if (!tir) take_refl = randf<R>(seed) < fr;
                if (take_refl) {
                    rd = refl;
                    ro = hp + nf * R(1e-3);
                    skip_sky = false;
                    prev_lambert = false;
                } else {
                    rd = normalize(refr);
                    ro = hp - nf * R(1e-3);
                }
            } else if (h.mat == MAT_WATER) {
                // the water surface (:6768-6906 without the FFT ocean and its
                // foam): the wave normal, Schlick with F0 0.02, reflection
                // (or total internal reflection) keeps the medium, refraction
                // crosses it; no NEE, the throughput unchanged
                const R wior = water_ior_of(h.ior);
                const V3<R> sn = water_normal(hp, h.n);
                const R ndi_s = dot(sn, rd);
                const V3<R> nf_s = (ndi_s < R(0)) ? sn : -sn;
                const R ia = (ndi_s < R(0)) ? R(1) : wior;
                const R ib = (ndi_s < R(0)) ? wior : R(1);
                const R eta = ia / ib;
                const R ci = oabs(dot(nf_s, -rd));
                const R r0 = R(0.02);
                const R m = R(1) - ci;
                const R fr = r0 + ((((R(1) - r0) * m) * m * m) * m) * m;
                const V3<R> refl = reflect(rd, nf_s);
                const V3<R> refr = refract(rd, nf_s, eta);
                bool take_refl = length(refr) < R(1e-3);
                if (!take_refl) take_refl = randf<R>(seed) < fr;
                if (take_refl) {
                    rd = refl;
                    ro = hp + nf_s * R(1e-3);
                    skip_sky = false;
                    prev_lambert = false;
                } else {
                    in_water = ndi_s < R(0);
                    rd = normalize(refr);
                    ro = hp - nf_s * R(1e-3);
                }
            }
            if (b >= 2u && !(F.flags & 1u)) {
                R p = clampv(maxc(T), R(0.05), R(1));
                if (randf<R>(seed) > p) break;
                T = T / p;
            }
        }
        return L;
    }

    void pixel(uint32_t x, uint32_t y, float* accum) const {
        uint32_t seed = pcg_hash(x * 1973u + y * 9277u + F.frame_index * 26699u);
        V3<R> fr{R(0), R(0), R(0)};
        uint32_t spp = std::max(1u, F.spp);
        for (uint32_t s = 0; s < spp; ++s) {
            V3<R> ro, rd;
            R time, lambda;
            camera(x, y, s, spp, seed, ro, rd, time, lambda);
            V3<R> L = path(ro, rd, seed, time, lambda);
            // the hero wavelength's sRGB response (PathTrace.slang:7693-7700)
            fr = fr + ((lambda >= R(0)) ? L * spectral_weight(lambda) : L);
        }
        fr = fr / R(spp);
        // NaN/Inf -> 0; negative channels -> 0 unless the frame is spectral
        // (PathTrace.slang:7722-7734: signed hero-wavelength values reach
        // the running mean, which stays unbiased; display clamps them)
        float o[3] = {float(fr.x), float(fr.y), float(fr.z)};
        for (int k = 0; k < 3; ++k) {
            if (std::isnan(o[k]) || std::isinf(o[k])) o[k] = 0.0f;
            else if (!F.spectral) o[k] = std::max(o[k], 0.0f);
        }
        float* a = accum + ((size_t)y * F.width + x) * 4;
        float prev[4] = {0, 0, 0, 0};
        if (F.reset_accum == 0u) {
            for (int k = 0; k < 4; ++k) prev[k] = a[k];
            bool bad = false;
            for (int k = 0; k < 4; ++k) bad = bad || std::isnan(prev[k]) || std::isinf(prev[k]);
            if (bad) prev[0] = prev[1] = prev[2] = prev[3] = 0.0f;
        }
        float ema = std::min(std::max(F.ema_alpha, 0.0f), 0.999f);
        if (prev[3] < 0.5f || ema <= 0.0f) {
            float nn = prev[3] + 1.0f;
            for (int k = 0; k < 3; ++k) a[k] = (prev[k] * prev[3] + o[k]) / nn;
            a[3] = nn;
        } else {
            for (int k = 0; k < 3; ++k) a[k] = prev[k] * ema + o[k] * (1.0f - ema);
            a[3] = std::min(prev[3] + 1.0f, 1.0f / std::max(1.0f - ema, 1e-3f));
        }
    }
};

template <class R>
void render_rows(const orc_scene& S, const orc_frame& F, int threads, uint32_t r0, uint32_t r1, uint32_t rs,
                 uint32_t cs, float* accum, Ctr& total) {
    std::atomic<uint32_t> next{0};
    uint32_t nrows = (r1 > r0) ? (r1 - r0 + rs - 1) / rs : 0;
    std::vector<Ctr> ctrs((size_t)threads);
    auto work = [&](int tid) {
        Ctr& c = ctrs[(size_t)tid];
        PT<R> pt{S, F, c};
        for (;;) {
            uint32_t k = next.fetch_add(1);
            if (k >= nrows) break;
            uint32_t y = r0 + k * rs;
            for (uint32_t x = 0; x < F.width; x += cs) pt.pixel(x, y, accum);
        }
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < threads; ++t) ts.emplace_back(work, t);
    work(0);
    for (auto& t : ts) t.join();
    for (auto& c : ctrs) {
        total.primary += c.primary; total.extension += c.extension; total.shadow += c.shadow;
        total.nodes += c.nodes; total.tests += c.tests;
    }
}

// d(frame value)/d(parameter) of pixel (x, y) by a PT<Dual> run; the
// sanitising step (NaN/Inf -> 0, max(0, .) outside spectral frames) is taken
// with derivative 1.
void pixel_tangent(const PT<Dual>& pt, const orc_frame& F, uint32_t x, uint32_t y, double dI[3], double I[3]) {
    uint32_t seed = pcg_hash(x * 1973u + y * 9277u + F.frame_index * 26699u);
    using R = Dual;
    V3<R> fr{R(0), R(0), R(0)};
    const uint32_t spp = std::max(1u, F.spp);
    for (uint32_t s = 0; s < spp; ++s) {
        V3<R> ro, rd;
        R time, lambda;
        pt.camera(x, y, s, spp, seed, ro, rd, time, lambda);
        V3<R> L = pt.path(ro, rd, seed, time, lambda);
        fr = fr + ((lambda >= R(0)) ? L * PT<Dual>::spectral_weight(lambda) : L);
    }
    fr = fr / R(spp);
    const Dual c[3] = {fr.x, fr.y, fr.z};
    for (int k = 0; k < 3; ++k) {
        const bool ok = std::isfinite(c[k].v) && std::isfinite(c[k].d);
        dI[k] = ok ? c[k].d : 0.0;
        I[k] = ok ? (F.spectral ? c[k].v : std::max(c[k].v, 0.0)) : 0.0;
    }
}

void fill_hit(orc_hit& o, const Hit<float>& h) {
    o.t = h.hit ? h.t : 0.0f;
    o.normal[0] = h.n.x; o.normal[1] = h.n.y; o.normal[2] = h.n.z;
    o.hit = h.hit ? h.kind : 0u;
    o.prim = h.hit ? h.prim : 0u;
    o.mat = h.hit ? h.mat : 0u;
    o.user_id = h.hit ? h.user_id : 0u;
}

}  // namespace

extern "C" {

int orc_scene_create(const orc_scene_desc* d, orc_scene** out) {
    if (!d || !out) return -1;
    auto* S = new orc_scene();
    S->prim_count = d->prim_count;
    if (d->prim_count) S->prims.assign(d->prims, d->prims + (size_t)d->prim_count * 28);
    std::vector<WorkItem> fw;
    for (uint32_t i = 0; i < d->prim_count; ++i) {
        const float* p = &S->prims[(size_t)i * 28];
        uint32_t type = (uint32_t)p[8];
        if (type == PRIM_PLANE) { S->planes.push_back(i); continue; }
        S->finite.push_back(i);
        WorkItem w{};
        for (int a = 0; a < 3; ++a) {
            // the sphere at its current and previous centre (motion blur moves
            // it between them), widened by 2^-20 of the coordinates so that
            // the rounding of the shutter-time lerp stays inside
            const float lo = std::min(p[a], p[12 + a]), hi = std::max(p[a], p[12 + a]);
            const float pad = (std::max(std::fabs(lo), std::fabs(hi)) + p[3]) * 9.5367431640625e-7f;
            w.c[a] = p[a];
            w.mn[a] = (lo - p[3]) - pad;
            w.mx[a] = (hi + p[3]) + pad;
        }
        w.id = i;
        fw.push_back(w);
    }
    S->use_abvh = S->finite.size() >= kAnalyticBvhThreshold;
    if (S->use_abvh) S->abvh.build(fw);
    if (d->index_count >= 3 && d->positions && d->indices) {
        S->pos.assign(d->positions, d->positions + (size_t)d->vertex_count * 3);
        S->idx.assign(d->indices, d->indices + (d->index_count / 3) * 3);
        S->tri_count = d->index_count / 3;
        for (uint32_t i = 0; i < S->idx.size(); ++i)
            if (S->idx[i] >= d->vertex_count) { delete S; return -1; }
        std::vector<WorkItem> tw(S->tri_count);
        for (uint32_t t = 0; t < S->tri_count; ++t) {
            const float* v0 = &S->pos[(size_t)S->idx[3 * t] * 3];
            const float* v1 = &S->pos[(size_t)S->idx[3 * t + 1] * 3];
            const float* v2 = &S->pos[(size_t)S->idx[3 * t + 2] * 3];
            WorkItem& w = tw[t];
            for (int a = 0; a < 3; ++a) {
                w.mn[a] = std::min(v0[a], std::min(v1[a], v2[a]));
                w.mx[a] = std::max(v0[a], std::max(v1[a], v2[a]));
                w.c[a] = (v0[a] + v1[a] + v2[a]) * (1.0f / 3.0f);
            }
            w.id = t;
        }
        S->tbvh.build(tw);
    }
    S->light_count = d->light_count;
    if (d->light_count) {
        S->lights.assign(d->lights, d->lights + (size_t)d->light_count * 20);
        S->ltree = build_light_tree(S->lights.data(), d->light_count);
    }
    if (d->env_rgb && d->env_width && d->env_height) {
        S->env_present = true;
        S->W = d->env_width;
        S->H = d->env_height;
        S->env.assign(d->env_rgb, d->env_rgb + (size_t)S->W * S->H * 3);
        S->env_intensity = d->env_intensity;
        if (d->env_marginal && d->env_conditional) {   // the engine's upload
            S->marg.assign(d->env_marginal, d->env_marginal + S->H);
            S->cond.assign(d->env_conditional, d->env_conditional + (size_t)S->W * S->H);
            S->env_total = float(d->env_total_luminance);
        } else {
            S->env_total = float(build_env_cdf(S->env.data(), S->W, S->H, S->marg, S->cond));
        }
    }
    if (d->sdf_clusters && d->sdf_cluster_count) {
        S->sdf_count = d->sdf_cluster_count;
        S->sdf.assign(d->sdf_clusters, d->sdf_clusters + (size_t)S->sdf_count * 112);
    }
    if (d->pbr_atlas && d->pbr_atlas_width == 384u && d->pbr_atlas_height >= 256u) {
        S->atlas_w = d->pbr_atlas_width;
        S->atlas_h = d->pbr_atlas_height;
        S->atlas_tiles = d->pbr_atlas_height / 256u;
        const size_t n = (size_t)S->atlas_w * S->atlas_h * 4;
        S->atlas.resize(n);
        for (size_t i = 0; i < n; ++i) S->atlas[i] = (float)d->pbr_atlas[i] * (1.0f / 255.0f);
    }
    if (d->mesh_uvs && d->vertex_count) S->uvs.assign(d->mesh_uvs, d->mesh_uvs + (size_t)d->vertex_count * 2);
    if (d->moon_rgba && d->moon_width && d->moon_height) {
        S->moon_w = d->moon_width;
        S->moon_h = d->moon_height;
        S->moon.assign(d->moon_rgba, d->moon_rgba + (size_t)S->moon_w * S->moon_h * 4);
    }
    *out = S;
    return 0;
}

void orc_scene_destroy(orc_scene* s) { delete s; }

int orc_sky(const orc_scene* s, const orc_frame* f, const float* dirs, uint32_t n, float* out) {
    if (!s || !f || (n && (!dirs || !out))) return -1;
    Ctr c;
    PT<float> pt{*s, *f, c};
    for (uint32_t i = 0; i < n; ++i) {
        V3<float> v = pt.sky(V3<float>{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]});
        out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
    }
    return 0;
}

void orc_sample_sun_disc(const float* dir, uint32_t seed0, uint32_t n, float* out) {
    orc_frame f{};
    orc_scene sc;
    Ctr c;
    PT<float> pt{sc, f, c};
    const V3<float> d{dir[0], dir[1], dir[2]};
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t seed = seed0 + i;
        V3<float> v = pt.sample_sun_disc(d, seed);
        out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
    }
}

int orc_render(const orc_scene* s, const orc_frame* f, int precision, int threads, uint32_t row_begin,
               uint32_t row_end, uint32_t row_step, uint32_t col_step, float* accum, orc_counts* counts) {
    if (!s || !f || !accum || row_step == 0 || col_step == 0) return -1;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    row_end = std::min(row_end, f->height);
    Ctr c;
    if (precision == 64) render_rows<double>(*s, *f, threads, row_begin, row_end, row_step, col_step, accum, c);
    else render_rows<float>(*s, *f, threads, row_begin, row_end, row_step, col_step, accum, c);
    if (counts) {
        counts->primary = c.primary; counts->extension = c.extension; counts->shadow = c.shadow;
        counts->node_visits = c.nodes; counts->prim_tests = c.tests;
    }
    return 0;
}

static int trace_impl(const orc_scene* s, const float* rays, uint32_t n, orc_hit* out, bool brute) {
    if (!s || (!rays && n) || (!out && n)) return -1;
    Ctr c;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + (size_t)i * 8;
        Hit<float> h = trace_scene<float>(*s, V3<float>{r[0], r[1], r[2]}, V3<float>{r[4], r[5], r[6]}, c, brute);
        fill_hit(out[i], h);
    }
    return 0;
}
int orc_trace(const orc_scene* s, const float* rays, uint32_t n, orc_hit* out) {
    return trace_impl(s, rays, n, out, false);
}
int orc_trace_bruteforce(const orc_scene* s, const float* rays, uint32_t n, orc_hit* out) {
    return trace_impl(s, rays, n, out, true);
}

void orc_sdf_eval(const float* cluster, const float* points, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const V3<float> p = ld3<float>(points + 3 * (size_t)i);
        const V3<float> g = normalize(Sdf<float>::grad(cluster, p));
        float* o = out + 4 * (size_t)i;
        o[0] = Sdf<float>::dist(cluster, p);
        o[1] = g.x; o[2] = g.y; o[3] = g.z;
    }
}

void orc_sdf_trace(const float* cluster, const float* rays, uint32_t n, uint32_t max_iters, float epsilon,
                   float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + (size_t)i * 8;
        const V3<float> o = ld3<float>(r), d = ld3<float>(r + 4);
        const V3<float> inv = safe_inv(d);
        float t = 0.0f;
        float* q = out + 4 * (size_t)i;
        if (Sdf<float>::march(cluster, o, d, inv, 1e-3f, 1e30f, max_iters, epsilon, t)) {
            const V3<float> g = normalize(Sdf<float>::grad(cluster, o + d * t));
            q[0] = t; q[1] = g.x; q[2] = g.y; q[3] = g.z;
        } else {
            q[0] = -1.0f; q[1] = q[2] = q[3] = 0.0f;
        }
    }
}

int orc_trace_at(const orc_scene* s, const float* rays, uint32_t n, float time, const float* shift_prev,
                 const float* shift_curr, int brute, orc_hit* out) {
    if (!s || (!rays && n) || (!out && n)) return -1;
    Ctr c;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + (size_t)i * 8;
        const V3<float> o = ld3<float>(r), d = ld3<float>(r + 4);
        V3<float> om = o;
        if (shift_prev && shift_curr) om = o - mix3(ld3<float>(shift_prev), ld3<float>(shift_curr), time);
        fill_hit(out[i], trace_scene<float>(*s, o, d, c, brute != 0, time, om));
    }
    return 0;
}

// G-buffer pass (PathTrace.slang:5463-5648): the pixel-centre ray (5464-5469),
// the hit's linear depth and surface fields or the sky sentinels (5495-5519),
// pixel motion by reprojection through the two view-projections (5524-5528),
// the MetalFX specular guidance fields (5611-5646).
int orc_gbuffer(const orc_scene* s, const orc_frame* f, const float* vp_curr, const float* vp_prev, float* out) {
    if (!s || !f || !vp_curr || !vp_prev || !out) return -1;
    using V = V3<float>;
    const float W = (float)f->width, H = (float)f->height;
    const V pos{f->pos_fovtan[0], f->pos_fovtan[1], f->pos_fovtan[2]};
    const float fovtan = f->pos_fovtan[3], aspect = f->fwd_aspect[3];
    const V fwd{f->fwd_aspect[0], f->fwd_aspect[1], f->fwd_aspect[2]};
    const V rt{f->right_xyz[0], f->right_xyz[1], f->right_xyz[2]};
    const V up{f->up_xyz[0], f->up_xyz[1], f->up_xyz[2]};
    const V fwdn = normalize(fwd);
    // mul(M, float4(p, 1)) for the column-major engine matrix
    auto project = [](const float* m, V p, float o[4]) {
        for (int i = 0; i < 4; ++i) o[i] = ((m[i] * p.x + m[4 + i] * p.y) + m[8 + i] * p.z) + m[12 + i];
    };
    auto pixel = [W, H](const float c[4], float& px, float& py) {
        const float cw = std::max(c[3], 1e-6f);
        px = ((c[0] / cw) * 0.5f + 0.5f) * W;
        py = ((-c[1] / cw) * 0.5f + 0.5f) * H;
    };
    Ctr c;
    for (uint32_t y = 0; y < f->height; ++y) {
        for (uint32_t x = 0; x < f->width; ++x) {
            const float ux = (((float)x + 0.5f) / W) * 2.0f - 1.0f;
            float uy = (((float)y + 0.5f) / H) * 2.0f - 1.0f;
            uy = -uy;
            const V d = normalize((fwd + rt * ((ux * aspect) * fovtan)) + up * (uy * fovtan));
            // shutter time 1: the mesh at its current translation (5470-5478)
            const V pm = f->mesh_shift_active ? pos - mix3(ld3<float>(f->mesh_shift_prev),
                                                            ld3<float>(f->mesh_shift_curr), 1.0f)
                                              : pos;
            const Hit<float> h = trace_scene<float>(*s, pos, d, c, false, 1.0f, pm);
            float* o = out + ((size_t)y * f->width + x) * 16u;
            for (int k = 0; k < 16; ++k) o[k] = 0.0f;
            V world;
            float depth, rough = 1.0f, dist = 0.0f;
            if (h.hit) {
                world = pos + d * h.t;
                depth = h.t * dot(d, fwdn);
                o[3] = h.n.x; o[4] = h.n.y; o[5] = h.n.z;
                o[6] = h.albedo.x; o[7] = h.albedo.y; o[8] = h.albedo.z;
                const float r = clampv(h.rough, 0.0f, 1.0f);
                if (h.mat == MAT_METAL) {
                    o[9] = h.albedo.x; o[10] = h.albedo.y; o[11] = h.albedo.z;
                } else if (h.mat == MAT_DIELECTRIC) {
                    o[9] = o[10] = o[11] = 0.04f;
                }
                rough = h.mat == MAT_LAMBERT ? 1.0f : r;
                dist = h.t * (1.0f - r);
            } else {
                world = pos + d * 1000.0f;
                depth = 1.0e10f;
            }
            float pc[4], pp[4], cx, cy, qx, qy;
            project(vp_curr, world, pc);
            project(vp_prev, world, pp);
            pixel(pc, cx, cy);
            pixel(pp, qx, qy);
            o[0] = depth;
            o[1] = qx - cx;
            o[2] = qy - cy;
            o[12] = rough;
            o[13] = dist;
        }
    }
    return 0;
}

int orc_tangent(orc_scene* s, const orc_frame* f, uint32_t prim, uint32_t param, const float* dimage, int threads,
                double* out_dloss, float* out_dimage) {
    if (!s || !f || !dimage || !out_dloss || prim >= s->prim_count || param > 6u) return -1;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    s->tangent_prim = prim;
    s->tangent_param = param;
    std::atomic<uint32_t> next{0};
    std::vector<double> part((size_t)threads, 0.0);
    auto work = [&](int tid) {
        Ctr c;
        PT<Dual> pt{*s, *f, c};
        double acc = 0.0;
        for (;;) {
            const uint32_t y = next.fetch_add(1);
            if (y >= f->height) break;
            for (uint32_t x = 0; x < f->width; ++x) {
                double dI[3], I[3];
                pixel_tangent(pt, *f, x, y, dI, I);
                const size_t p = ((size_t)y * f->width + x) * 4u;
                for (int k = 0; k < 3; ++k) acc += (double)dimage[p + k] * dI[k];
                if (out_dimage)
                    for (int k = 0; k < 3; ++k) out_dimage[p + k] = (float)dI[k];
            }
        }
        part[(size_t)tid] = acc;
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < threads; ++t) ts.emplace_back(work, t);
    work(0);
    for (auto& t : ts) t.join();
    double total = 0.0;
    for (double v : part) total += v;
    *out_dloss = total;
    s->tangent_prim = 0xFFFFFFFFu;
    return 0;
}

int orc_render_f64_image(orc_scene* s, const orc_frame* f, int threads, double* out_rgb) {
    if (!s || !f || !out_rgb) return -1;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    s->tangent_prim = 0xFFFFFFFFu;
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        Ctr c;
        PT<Dual> pt{*s, *f, c};
        for (;;) {
            const uint32_t y = next.fetch_add(1);
            if (y >= f->height) break;
            for (uint32_t x = 0; x < f->width; ++x) {
                double dI[3], I[3];
                pixel_tangent(pt, *f, x, y, dI, I);
                for (int k = 0; k < 3; ++k) out_rgb[((size_t)y * f->width + x) * 3u + k] = I[k];
            }
        }
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < threads; ++t) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
    return 0;
}

int orc_pbr_sample(const orc_scene* s, uint32_t tile, const float* uvl, uint32_t n, float* out) {
    if (!s || s->atlas.empty() || tile >= s->atlas_tiles) return 1;
    orc_frame f{};
    Ctr c;
    PT<float> pt{*s, f, c};
    for (uint32_t i = 0; i < n; ++i) {
        const auto r = pt.tile_linear(tile, uvl[3 * i], uvl[3 * i + 1], uvl[3 * i + 2]);
        out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
    }
    return 0;
}

int orc_pbr_surface(const orc_scene* s, const orc_frame* f, const float* rays, uint32_t n, float cone_width, float* out) {
    if (!s || !f) return 1;
    Ctr c;
    PT<float> pt{*s, *f, c};
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        const V3<float> ro{r[0], r[1], r[2]}, rd{r[4], r[5], r[6]};
        Hit<float> h = pt.trace(ro, rd, 1.0f);
        float* o = out + 16 * i;
        for (int k = 0; k < 16; ++k) o[k] = 0.0f;
        if (!h.hit) continue;
        uint32_t tiles[4];
        float u, v, scale, lod = 0.0f;
        pt.pbr_uv(h, ro, rd, 1.0f, tiles, u, v, scale);
        if (!s->atlas.empty()) pt.apply_pbr(h, ro, rd, 1.0f, cone_width, &lod);
        const float vals[16] = {u, v, scale, lod, h.albedo.x, h.albedo.y, h.albedo.z, h.rough, (float)h.mat,
                                h.n.x, h.n.y, h.n.z, h.t, (float)h.kind, 0.0f, 0.0f};
        for (int k = 0; k < 16; ++k) o[k] = vals[k];
    }
    return 0;
}

void orc_pbr_normal_map(const float* n, const float* ts, uint32_t count, float* out) {
    for (uint32_t i = 0; i < count; ++i) {
        const auto r = PT<float>::normal_map(V3<float>{n[3 * i], n[3 * i + 1], n[3 * i + 2]},
                                             V3<float>{ts[3 * i], ts[3 * i + 1], ts[3 * i + 2]});
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
}

void orc_pbr_srgb(const float* c, uint32_t count, float* out) {
    for (uint32_t i = 0; i < count; ++i) out[i] = PT<float>::srgb_to_linear(c[i]);
}

void orc_value_noise(const float* p, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = PT<float>::value_noise(V3<float>{p[3 * i], p[3 * i + 1], p[3 * i + 2]});
}

int orc_water_normal(const orc_frame* f, const float* p, const float* nm, uint32_t n, float* out) {
    if (!f) return 1;
    orc_scene s;
    Ctr c;
    PT<float> pt{s, *f, c};
    for (uint32_t i = 0; i < n; ++i) {
        const auto r = pt.water_normal(V3<float>{p[3 * i], p[3 * i + 1], p[3 * i + 2]},
                                       V3<float>{nm[3 * i], nm[3 * i + 1], nm[3 * i + 2]});
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
    return 0;
}

double orc_build_env_cdf(const float* rgb, uint32_t W, uint32_t H, float* om, float* oc) {
    std::vector<float> m, c;
    double t = build_env_cdf(rgb, W, H, m, c);
    std::copy(m.begin(), m.end(), om);
    std::copy(c.begin(), c.end(), oc);
    return t;
}

uint32_t orc_build_light_tree(const float* lights, uint32_t n, float* out_nodes) {
    auto t = build_light_tree(lights, n);
    if (!t.empty()) std::memcpy(out_nodes, t.data(), t.size() * sizeof(LightTreeNode));
    return (uint32_t)t.size();
}

uint32_t orc_build_tri_bvh(const float* pos, uint32_t vc, const uint32_t* idx, uint32_t ic, uint32_t* on,
                           uint32_t* op) {
    std::vector<WorkItem> tw;
    uint32_t tc = ic / 3;
    for (uint32_t t = 0; t < tc; ++t) {
        uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
        if (i0 >= vc || i1 >= vc || i2 >= vc) continue;   // TriangleBvh.cpp:49
        const float *v0 = pos + 3 * (size_t)i0, *v1 = pos + 3 * (size_t)i1, *v2 = pos + 3 * (size_t)i2;
        WorkItem w{};
        for (int a = 0; a < 3; ++a) {
            w.mn[a] = std::min(v0[a], std::min(v1[a], v2[a]));
            w.mx[a] = std::max(v0[a], std::max(v1[a], v2[a]));
            w.c[a] = (v0[a] + v1[a] + v2[a]) * (1.0f / 3.0f);
        }
        w.id = t;
        tw.push_back(w);
    }
    MedianBvh b;
    b.build(tw);
    if (!b.nodes.empty()) std::memcpy(on, b.nodes.data(), b.nodes.size() * sizeof(BNode));
    if (!b.perm.empty()) std::memcpy(op, b.perm.data(), b.perm.size() * 4);
    return (uint32_t)b.nodes.size();
}

int orc_env_sample(const orc_scene* s, const float* xi, uint32_t n, float* out) {
    if (!s) return -1;
    orc_frame f{};
    Ctr c;
    PT<float> pt{*s, f, c};
    for (uint32_t i = 0; i < n; ++i) {
        auto e = pt.sample_env(xi[2 * i], xi[2 * i + 1]);
        float* o = out + 8 * (size_t)i;
        o[0] = e.dir.x; o[1] = e.dir.y; o[2] = e.dir.z;
        o[3] = e.color.x; o[4] = e.color.y; o[5] = e.color.z;
        o[6] = e.pdf; o[7] = 0.0f;
    }
    return 0;
}

int orc_env_pdf(const orc_scene* s, const float* dirs, uint32_t n, float* out) {
    if (!s) return -1;
    orc_frame f{};
    Ctr c;
    PT<float> pt{*s, f, c};
    for (uint32_t i = 0; i < n; ++i) out[i] = pt.pdf_env(V3<float>{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]});
    return 0;
}

// Exhaustive pick PMF of the light-tree descent at (p, nf): the product of
// the branch probabilities of pickLightFromTree along each root-to-leaf path.
int orc_light_pmf(const orc_scene* s, const float* p, const float* nf, float* pmf) {
    if (!s || s->ltree.empty()) return -1;
    orc_frame f{};
    Ctr c;
    PT<float> pt{*s, f, c};
    V3<float> P{p[0], p[1], p[2]}, N{nf[0], nf[1], nf[2]};
    for (uint32_t i = 0; i < s->light_count; ++i) pmf[i] = 0.0f;
    std::vector<std::pair<uint32_t, double>> st{{0u, 1.0}};
    while (!st.empty()) {
        auto [idx, pr] = st.back();
        st.pop_back();
        const LightTreeNode& nd = s->ltree[idx];
        if (nd.count > 0u) { pmf[nd.left_first] += float(pr); continue; }
        float il = pt.importance(s->ltree[nd.left_first], P, N);
        float ir = pt.importance(s->ltree[nd.left_first + 1], P, N);
        float tot = il + ir;
        float pl = (tot > 1e-30f) ? (il / tot) : 0.5f;
        st.push_back({nd.left_first, pr * pl});
        st.push_back({nd.left_first + 1, pr * (1.0f - pl)});
    }
    return 0;
}

void orc_math_f32(int fn, const float* a, const float* b, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = MathF::sin(a[i]); break;
            case 1: out[i] = MathF::cos(a[i]); break;
            case 2: out[i] = MathF::acos(a[i]); break;
            case 3: out[i] = MathF::atan2(a[i], b[i]); break;
            case 4: out[i] = MathF::exp(a[i]); break;
            case 5: out[i] = PT<float>::spectral_weight(a[i]).x; break;
            case 6: out[i] = PT<float>::spectral_weight(a[i]).y; break;
            case 7: out[i] = PT<float>::spectral_weight(a[i]).z; break;
            case 8: out[i] = PT<float>::cauchy_ior(b[i], 0.0042f, a[i]); break;
            case 9: out[i] = MathF::log(a[i]); break;
            case 10: out[i] = MathF::pow(a[i], b[i]); break;
            default: out[i] = 0.0f;
        }
    }
}

int orc_lens_sample(const float* u, const float* uv, uint32_t n, int blades, float rot, float cat_eye, float* out) {
    if (!u || !out || (cat_eye > 0.0f && !uv)) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        float x, y;
        if (blades >= 3) PT<float>::polygon_sample(u[2 * i], u[2 * i + 1], blades, rot, x, y);
        else PT<float>::concentric_disk(u[2 * i], u[2 * i + 1], x, y);
        if (cat_eye > 0.0f) PT<float>::cat_eye(x, y, uv[2 * i], uv[2 * i + 1], cat_eye);
        out[2 * i] = x;
        out[2 * i + 1] = y;
    }
    return 0;
}

}  // extern "C"
