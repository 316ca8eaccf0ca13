// Device shading helpers of the path tracer: sky / env-map sampling, light
// tree descent, analytic light sampling, cosine hemisphere sampling.  Each
// function names the reference shader code it evaluates.
#pragma once
#include "dmt_dev.h"

namespace dmt {

constexpr float kPi = 3.14159265358979f;

struct FrameArgs {
    f3 pos, fwd, right, up;
    float fovtan, aspect;
    uint32_t width, height;
    uint32_t frame_index;
    uint32_t spp;
    uint32_t max_bounces;
    uint32_t reset_accum;
    uint32_t mis_enabled;
    uint32_t light_tree_enabled;
    uint32_t refract_bounces;
    float firefly_clamp;
    float ema_alpha;
    uint32_t rr_enabled;      // Russian roulette from bounce 2 on
    float cauchy_c;           // Cauchy C (um^2) of the spectral dielectric index
};

// Camera-sample features of a frame (all off: the plain camera ray).
struct FeatureArgs {
    float aperture, focal, rot, cat_eye;   // thin lens (dof_params, dof_bokeh_params.x)
    int blades;                            // < 3: round aperture
    uint32_t motion_blur;                  // draw a shutter time per sample
    uint32_t spectral;                     // draw a hero wavelength per sample
    uint32_t pad;
};

// clampIndirect (PathTrace.slang:3367-3371).
template <class T> __device__ __forceinline__ V3<T> clamp_indirect(const FrameArgs& F, V3<T> c) {
    if (F.firefly_clamp <= 0.0f) return c;
    T m = maxc(c);
    return (m > F.firefly_clamp) ? c * (F.firefly_clamp / m) : c;
}

// sampleCosineHemisphere (PathTraceMath.slang:30-42).
template <class T> __device__ __forceinline__ V3<T> cos_sample(V3<T> n, uint32_t& seed) {
    float u1 = randf(seed);
    float u2 = randf(seed);
    float r = sqrtf(u1);
    float phi = 6.28318530718f * u2;
    float sp, cp;
    cephes_sincos(phi, &sp, &cp);
    float x = r * cp;
    float y = r * sp;
    float z = sqrtf(fmaxf(0.0f, 1.0f - u1));
    V3<T> ax = to<T>(fabs_(n.x) > 0.9f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f});
    V3<T> t = normalize(cross(n, ax));
    V3<T> b = cross(n, t);
    return normalize((t * x + b * y) + n * z);
}

__device__ __forceinline__ f3 env_texel(const SceneView& S, uint32_t x, uint32_t y) {
    return xyz(S.env[(size_t)y * S.env_w + x]);
}

// skyColor (PathTrace.slang:3373-3431): bilinear lat-long lookup; without an
// env map the base gradient (sun and moon below the horizon).
template <class T> __device__ __forceinline__ V3<T> sky_color(const SceneView& S, V3<T> rd) {
    if (!S.env_present) {
        T t = clampf(0.5f * (rd.y + 1.0f), 0.0f, 1.0f);
        return lerp3(to<T>(f3{0.5f, 0.55f, 0.62f}), to<T>(f3{1.0f, 1.0f, 1.0f}), t);
    }
    const int W = (int)S.env_w, H = (int)S.env_h;
    T u = c_atan2(rd.z, rd.x) * (1.0f / 6.28318530718f) + 0.5f;
    T v = c_acos(clampf(rd.y, -1.0f, 1.0f)) * (1.0f / 3.14159265359f);
    T fx = u * (float)W - 0.5f, fy = v * (float)H - 0.5f;
    int ix = (int)floorf(primal(fx)), iy = (int)floorf(primal(fy));
    T ax = fx - (float)ix, ay = fy - (float)iy;
    int x0 = (ix % W + W) % W;
    int x1 = (x0 + 1) % W;
    int y0 = min(max(iy, 0), H - 1);
    int y1 = min(max(iy + 1, 0), H - 1);
    V3<T> c00 = to<T>(env_texel(S, x0, y0)), c10 = to<T>(env_texel(S, x1, y0));
    V3<T> c01 = to<T>(env_texel(S, x0, y1)), c11 = to<T>(env_texel(S, x1, y1));
    V3<T> s = lerp3(lerp3(c00, c10, ax), lerp3(c01, c11, ax), ay);
    return s * S.env_intensity;
}

// cdfSearch / cdfSearchMarginal (PathTrace.slang:3619-3636).
__device__ __forceinline__ uint32_t cdf_search(const float* cdf, uint32_t count, float u) {
    uint32_t lo = 0, hi = count;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return min(lo, count - 1u);
}

struct EnvSample {
    f3 dir, color;
    float pdf;
};

// sampleEnvMap (PathTrace.slang:3638-3670).
__device__ __forceinline__ EnvSample sample_env(const SceneView& S, float xi0, float xi1) {
    EnvSample s{f3{0.0f, 1.0f, 0.0f}, f3{0.0f, 0.0f, 0.0f}, 0.0f};
    if (!S.env_present || !(S.env_total > 0.0f)) return s;
    const uint32_t W = S.env_w, H = S.env_h;
    uint32_t vi = cdf_search(S.env_marg, H, xi1);
    uint32_t ui = cdf_search(S.env_cond + (size_t)vi * W, W, xi0);
    float u = ((float)ui + 0.5f) / (float)W;
    float v = ((float)vi + 0.5f) / (float)H;
    float theta = v * kPi;
    float phi = u * 6.28318530717958f - kPi;
    float st, ct, sp, cp;
    cephes_sincos(theta, &st, &ct);
    cephes_sincos(phi, &sp, &cp);
    s.dir = f3{st * cp, ct, st * sp};
    f3 raw = env_texel(S, ui, vi);
    s.color = raw * S.env_intensity;
    float lum = (0.2126f * raw.x + 0.7152f * raw.y) + 0.0722f * raw.z;
    if (st > 1e-6f) s.pdf = (lum * (float)(W * H)) / (((2.0f * kPi) * kPi) * S.env_total);
    return s;
}

// pdfEnvMapDir (PathTrace.slang:3710-3727).
__device__ __forceinline__ float pdf_env(const SceneView& S, f3 d) {
    if (!S.env_present || !(S.env_total > 0.0f)) return 0.0f;
    float th = c_acos(clampf(d.y, -1.0f, 1.0f));
    float st = c_sin(th);
    if (st < 1e-6f) return 0.0f;
    float ph = c_atan2(d.z, d.x);
    const uint32_t W = S.env_w, H = S.env_h;
    float u = (ph + kPi) * (1.0f / 6.28318530717958f);
    float v = th * (1.0f / kPi);
    uint32_t ui = min((uint32_t)(u * (float)W), W - 1u);
    uint32_t vi = min((uint32_t)(v * (float)H), H - 1u);
    f3 raw = env_texel(S, ui, vi);
    float lum = (0.2126f * raw.x + 0.7152f * raw.y) + 0.0722f * raw.z;
    return (lum * (float)(W * H)) / (((2.0f * kPi) * kPi) * S.env_total);
}

// The env pdf is piecewise constant in the direction (one texel's
// luminance): its derivative is zero, the generic form evaluates the float one.
template <class T> __device__ __forceinline__ T pdf_env_of(const SceneView& S, V3<T> d) {
    return T(pdf_env(S, f3{primal(d.x), primal(d.y), primal(d.z)}));
}

// misBalance (PathTrace.slang:3733-3735).
template <class T> __device__ __forceinline__ T mis_balance(T a, T b) { return a / fmax_(a + b, T(1e-12f)); }

// lightTreeImportance (PathTrace.slang:3929-3966).
template <class T> __device__ __forceinline__ T lt_importance(const float4* nd, V3<T> p, V3<T> nf) {
    const float4 a = nd[0], b = nd[1], c = nd[2], e = nd[3];
    f3 mn = xyz(a), mx = xyz(b);
    f3 ce = (mn + mx) * 0.5f;
    V3<T> to_c = to<T>(ce) - p;
    T d2 = fmax_(dot(to_c, to_c), T(1e-12f));
    T d = sqrt_(d2);
    V3<T> dir = to_c / d;
    f3 ext = (mx - mn) * 0.5f;
    float rad2 = fmaxf(dot(ext, ext), 1e-12f);
    T fall = 1.0f / fmax_(d2, T(rad2));
    T orient = T(1.0f);
    if (c.w > -0.999f) {
        T ce2 = -dot(to<T>(xyz(c)), dir);
        T t = (ce2 - (c.w - 0.2f)) / 0.2f;
        orient = clampf(t, 0.0f, 1.0f);
    }
    T recv = fmax_(T(0.0f), dot(nf, dir));
    return fmax_(((e.x * fall) * orient) * recv, T(1e-8f));
}

template <class T> struct LightSampleT {
    V3<T> dir, Le;
    T pdf, dist;
};
using LightSample = LightSampleT<float>;

// pickLightFromTree + sampleAnalyticLight (PathTrace.slang:3979-4179).
template <class T>
__device__ __forceinline__ LightSampleT<T> sample_light(const SceneView& S, const FrameArgs& F, V3<T> p, V3<T> nf,
                                                       uint32_t& seed) {
    LightSampleT<T> s{to<T>(f3{0.0f, 1.0f, 0.0f}), to<T>(f3{0.0f, 0.0f, 0.0f}), T(0.0f), T(1e30f)};
    const uint32_t n = S.light_count;
    if (n == 0u) return s;
    uint32_t pick;
    T pick_pdf;
    if (F.light_tree_enabled != 0u && S.ltree_count > 0u) {
        uint32_t idx = 0u, lid = 0u;
        pick_pdf = T(1.0f);
        for (uint32_t dd = 0; dd < 32u; ++dd) {
            const float4* nd = S.ltree + (size_t)idx * 4u;
            const float4 v0 = nd[0], v1 = nd[1];
            if (__float_as_uint(v1.w) > 0u) { lid = __float_as_uint(v0.w); break; }
            const uint32_t l = __float_as_uint(v0.w), r = l + 1u;
            T il = lt_importance(S.ltree + (size_t)l * 4u, p, nf);
            T ir = lt_importance(S.ltree + (size_t)r * 4u, p, nf);
            T tot = il + ir;
            T pl = (tot > 1e-30f) ? (il / tot) : T(0.5f);
            float u = randf(seed);
            if (u < pl) { idx = l; pick_pdf = pick_pdf * pl; }
            else { idx = r; pick_pdf = pick_pdf * (1.0f - pl); }
        }
        pick = min(lid, n - 1u);
    } else {
        float u = randf(seed);
        pick = min((uint32_t)(u * (float)n), n - 1u);
        pick_pdf = T(1.0f / (float)n);
    }
    const float4* L = S.lights + (size_t)pick * 5u;
    const float4 l0 = L[0], l1 = L[1], l2 = L[2], l3 = L[3], l4 = L[4];
    const uint32_t type = (uint32_t)(l1.w + 0.5f);
    const f3 pos = xyz(l0);
    const f3 I = xyz(l1);
    if (type == 0u) {   // point
        V3<T> tl = to<T>(pos) - p;
        T d2 = fmax_(dot(tl, tl), T(1e-12f));
        T d = sqrt_(d2);
        s.dir = tl / d; s.Le = to<T>(I) / d2; s.pdf = pick_pdf; s.dist = d;
        return s;
    }
    if (type == 1u) {   // spot
        V3<T> tl = to<T>(pos) - p;
        T d2 = fmax_(dot(tl, tl), T(1e-12f));
        T d = sqrt_(d2);
        V3<T> tld = tl / d;
        f3 axis = normalize(quat_rotate(l4, xyz(l2)));
        T cd = dot(-tld, to<T>(axis));
        float ci = l3.x, co = l2.w;
        T fo = T(0.0f);
        if (cd >= ci) fo = T(1.0f);
        else if (cd > co) fo = (cd - co) / fmaxf(ci - co, 1e-6f);
        s.dir = tld; s.Le = to<T>(I) * (fo / d2); s.pdf = pick_pdf; s.dist = d;
        return s;
    }
    if (type == 2u) {   // sphere
        float x0 = randf(seed);
        float x1 = randf(seed);
        float z = 1.0f - 2.0f * x0;
        float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float phi = 6.28318530717958f * x1;
        float sp, cp;
        cephes_sincos(phi, &sp, &cp);
        f3 nl{r * cp, r * sp, z};
        f3 lp = pos + nl * l0.w;
        V3<T> tl = to<T>(lp) - p;
        T d2 = fmax_(dot(tl, tl), T(1e-12f));
        T d = sqrt_(d2);
        V3<T> wi = tl / d;
        T cl = fmax_(T(0.0f), dot(to<T>(nl), -wi));
        if (cl <= 0.0f) return s;
        float area = ((4.0f * kPi) * l0.w) * l0.w;
        T pdf_sa = d2 / fmax_(area * cl, T(1e-12f));
        s.dir = wi; s.Le = to<T>(I); s.pdf = pick_pdf * pdf_sa; s.dist = d;
        return s;
    }
    if (type == 4u) {   // emissive triangle (a = v0, b = v2, c = v3), uniform in area, one-sided
        float x0 = randf(seed);
        float x1 = randf(seed);
        float su = sqrtf(x0);
        float b0 = 1.0f - su;
        float b1 = x1 * su;
        const f3 a = pos, b = xyz(l2), c = xyz(l3);
        f3 lp = (a * b0 + b * b1) + c * ((1.0f - b0) - b1);
        f3 cr = cross(b - a, c - a);
        float area = 0.5f * length(cr);
        f3 nt = normalize(cr);
        V3<T> tl = to<T>(lp) - p;
        T d2 = fmax_(dot(tl, tl), T(1e-12f));
        T d = sqrt_(d2);
        V3<T> wi = tl / d;
        T cl = dot(to<T>(nt), -wi);
        if (cl <= 0.0f) return s;
        T pdf_sa = d2 / fmax_(area * cl, T(1e-12f));
        s.dir = wi; s.Le = to<T>(I); s.pdf = pick_pdf * pdf_sa; s.dist = d;
        return s;
    }
    if (type == 3u) {   // quad
        float x0 = randf(seed) * 2.0f - 1.0f;
        float x1 = randf(seed) * 2.0f - 1.0f;
        f3 uv = quat_rotate(l4, xyz(l3));
        float ul = length(uv);
        if (ul < 1e-6f) return s;
        f3 nq = normalize(quat_rotate(l4, xyz(l2)));
        f3 va = normalize(cross(nq, uv));
        f3 lp = (pos + uv * x0) + va * (l3.w * x1);
        V3<T> tl = to<T>(lp) - p;
        T d2 = fmax_(dot(tl, tl), T(1e-12f));
        T d = sqrt_(d2);
        V3<T> wi = tl / d;
        T cl = dot(to<T>(nq), -wi);
        if (cl <= 0.0f) return s;
        float area = (4.0f * ul) * l3.w;
        T pdf_sa = d2 / fmax_(area * cl, T(1e-12f));
        s.dir = wi; s.Le = to<T>(I); s.pdf = pick_pdf * pdf_sa; s.dist = d;
        return s;
    }
    return s;
}

}  // namespace dmt
