// One path vertex of the reference path tracer, shared by the wavefront
// kernels and the persistent (megakernel) form so both evaluate identical
// arithmetic.
//
//   camera_ray      PathTrace.slang:5729-5778 (jitter :5733)
//   shade_vertex    miss / MIS sky (:6034-6153), emission (:6193-6199),
//                   Lambert (:6320-6682), metal (:6683-6708),
//                   dielectric (:6709-6767), Russian roulette (:6908-6913)
//   transmittance   :5089-5196
//   shadow_value    the light-NEE (:6353-6364) and env-NEE (:6466-6492)
//                   contribution once the transmittance is known
#pragma once
#include "shade.h"

namespace dmt {

// Camera ray of pixel (x, y), drawing the two jitter numbers from seed.
__device__ __forceinline__ f3 camera_ray(const FrameArgs& F, uint32_t x, uint32_t y, uint32_t& seed) {
    float jx = randf(seed);
    float jy = randf(seed);
    jx = jx - 0.5f;
    jy = jy - 0.5f;
    float ux = ((((float)x + 0.5f) + jx) / (float)F.width) * 2.0f - 1.0f;
    float uy = ((((float)y + 0.5f) + jy) / (float)F.height) * 2.0f - 1.0f;
    uy = -uy;
    return normalize((F.fwd + F.right * ((ux * F.aspect) * F.fovtan)) + F.up * (uy * F.fovtan));
}

__device__ __forceinline__ uint32_t pixel_seed(const FrameArgs& F, uint32_t x, uint32_t y, uint32_t frame) {
    return pcg_hash(x * 1973u + y * 9277u + frame * 26699u);
}

// A shadow query and the factors of its contribution.  kind 0 = light NEE,
// 1 = env NEE; pdf < 0 marks the no-env ambient term (added without tracing).
template <class T> struct ShadowRecT {
    V3<T> o, d;
    T t_max;
    V3<T> pre;     // throughput * Le (light) or throughput * env colour
    T ndl;
    V3<T> brdf;
    T pdf;
    T w;           // MIS weight (env)
    bool clamp;    // clampIndirect applies (light NEE at b > 0; env always)
};
using ShadowRec = ShadowRecT<float>;

template <class T> struct PathVertexStateT {
    V3<T> T_, L;
    T prev_pdf;
    bool skip_sky, prev_lambert;
    uint32_t seed;
};

using PathVertexState = PathVertexStateT<float>;

template <class T> struct VertexResultT {
    bool alive;                  // the path scatters (before the depth limit)
    V3<T> no, nd;                // next ray
};
using VertexResult = VertexResultT<float>;

// Observation points of shade_vertex for the adjoint (path replay) sinks;
// the rendering sinks derive from this and ignore them.
struct NoGradientHooks {
    // A contribution added to L at this vertex: v before clampIndirect, c after.
    template <class T> __device__ __forceinline__ void on_radiance(V3<T> v, V3<T> c, bool may_clamp) {}
    // The emission term of hit surface sf (called before on_radiance for it).
    template <class T>
    __device__ __forceinline__ void on_emission(const SurfaceT<T>& sf, V3<T> T_, V3<T> v, V3<T> c, bool may_clamp) {}
    // The throughput is about to be multiplied by sf.albedo (Lambert, metal);
    // called before the vertex's own NEE terms, which also carry sf.albedo.
    template <class T> __device__ __forceinline__ void on_albedo_vertex(const SurfaceT<T>& sf) {}
    // The surface of bounce b before it is shaded (the tangent pass seeds a
    // roughness derivative here).
    template <class T> __device__ __forceinline__ void on_surface(uint32_t b, SurfaceT<T>& sf) {}
};

// Shades the vertex at bounce b for the ray (o, d) with closest hit h.
// Shadow queries go to sink.emit(kind, record) as soon as they are formed
// (light NEE before env NEE, the order their contributions are summed).
template <class T, class Sink>
__device__ __forceinline__ void shade_vertex(const SceneView& S, const FrameArgs& F, uint32_t b, V3<T> o, V3<T> d,
                                             HitRec h, PathVertexStateT<T>& P, VertexResultT<T>& R, Sink& sink) {
    const bool mis_on = F.mis_enabled != 0u;
    R.alive = false;
    R.no = to<T>(f3{0.0f, 0.0f, 0.0f});
    R.nd = to<T>(f3{0.0f, 1.0f, 0.0f});
    if (h.code == kMiss) {
        if (mis_on && P.prev_lambert) {
            // transmittance(ro, rd, 1e30) would re-trace this missed ray and
            // return exactly 1 (PathTrace.slang:6081-6084).
            T pe = pdf_env_of(S, d);
            T w = mis_balance(P.prev_pdf, pe);
            const V3<T> v = (P.T_ * sky_color(S, d)) * w;
            const V3<T> c = clamp_indirect(F, v);
            P.L = P.L + c;
            sink.on_radiance(v, c, true);
        } else if (!P.skip_sky) {
            V3<T> v = P.T_ * sky_color(S, d);
            const V3<T> c = (b == 0u) ? v : clamp_indirect(F, v);
            P.L = P.L + c;
            sink.on_radiance(v, c, b != 0u);
        }
        return;
    }
    SurfaceT<T> sf = surface_of(S, o, d, h);
    sink.on_surface(b, sf);
    const V3<T> hp = o + d * sf.t;
    const T ndi = dot(sf.n, d);
    const V3<T> nf = (ndi < 0.0f) ? sf.n : -sf.n;
    if (any_pos(sf.emission)) {
        V3<T> v = P.T_ * to<T>(sf.emission);
        const V3<T> c = (b == 0u) ? v : clamp_indirect(F, v);
        P.L = P.L + c;
        sink.on_emission(sf, P.T_, v, c, b != 0u);
        sink.on_radiance(v, c, b != 0u);
    }
    if (sf.mat == 0u) {
        sink.on_albedo_vertex(sf);
        if (S.light_count > 0u) {
            LightSampleT<T> ls = sample_light(S, F, hp, nf, P.seed);
            T ndl = dot(nf, ls.dir);
            if (ls.pdf > 0.0f && ndl > 0.0f) {
                ShadowRecT<T> r;
                r.o = hp + nf * 1e-3f;
                r.d = ls.dir;
                r.t_max = fmax_(ls.dist - 1e-3f, T(0.0f));
                r.pre = P.T_ * ls.Le;
                r.ndl = ndl;
                r.brdf = to<T>(sf.albedo / kPi);
                r.pdf = ls.pdf;
                r.w = T(1.0f);
                r.clamp = b > 0u;
                sink.emit(0, r, P);
            }
        }
        if (S.env_present) {
            float x0 = randf(P.seed);
            float x1 = randf(P.seed);
            EnvSample e = sample_env(S, x0, x1);
            T ndl = dot(nf, to<T>(e.dir));
            if (e.pdf > 0.0f && ndl > 0.0f) {
                T pb = ndl * (1.0f / kPi);
                bool fire = mis_on && (b + 1u) < F.max_bounces;
                ShadowRecT<T> r;
                r.o = hp + nf * 1e-3f;
                r.d = to<T>(e.dir);
                r.t_max = T(1e30f);
                r.pre = P.T_ * to<T>(e.color);
                r.ndl = ndl;
                r.brdf = to<T>(sf.albedo / kPi);
                r.pdf = T(e.pdf);
                r.w = fire ? mis_balance(T(e.pdf), pb) : T(1.0f);
                r.clamp = true;
                sink.emit(1, r, P);
            }
            P.skip_sky = !mis_on;
        } else {
            // No env map: ambient albedo * skyColorBase(nf) (PathTrace.slang:6667),
            // added after the light NEE term.
            T tt = clampf(0.5f * (nf.y + 1.0f), 0.0f, 1.0f);
            ShadowRecT<T> r;
            r.o = hp;
            r.d = to<T>(f3{0.0f, 1.0f, 0.0f});
            r.t_max = T(0.0f);
            r.pre = (P.T_ * to<T>(sf.albedo)) * lerp3(to<T>(f3{0.5f, 0.55f, 0.62f}), to<T>(f3{1.0f, 1.0f, 1.0f}), tt);
            r.ndl = T(0.0f);
            r.brdf = to<T>(f3{0.0f, 0.0f, 0.0f});
            r.pdf = T(-1.0f);
            r.w = T(0.0f);
            r.clamp = false;
            sink.emit(1, r, P);
            P.skip_sky = true;
        }
        P.T_ = P.T_ * to<T>(sf.albedo);
        R.no = hp + nf * 1e-3f;
        R.nd = cos_sample(nf, P.seed);
        if (mis_on && S.env_present) {
            P.prev_pdf = fmax_(T(0.0f), dot(nf, R.nd)) * (1.0f / kPi);
            P.prev_lambert = true;
        } else {
            P.prev_lambert = false;
        }
        R.alive = true;
    } else if (sf.mat == 1u) {
        sink.on_albedo_vertex(sf);
        V3<T> r = reflect(d, nf);
        V3<T> cs = cos_sample(nf, P.seed);
        r = normalize(r + (cs - nf) * sf.rough);
        P.T_ = P.T_ * to<T>(sf.albedo);
        R.no = hp + nf * 1e-3f;
        R.nd = r;
        P.skip_sky = false;
        P.prev_lambert = false;
        R.alive = true;
    } else {
        float ia = (ndi < 0.0f) ? 1.0f : sf.ior;
        float ib = (ndi < 0.0f) ? sf.ior : 1.0f;
        float eta = ia / ib;
        T ci = fabs_(dot(nf, -d));
        float r0 = (ia - ib) / (ia + ib);
        r0 = r0 * r0;
        T m = 1.0f - ci;
        T fr = r0 + ((((1.0f - r0) * m) * m * m) * m) * m;
        V3<T> refl = reflect(d, nf);
        V3<T> refr = refract(d, nf, eta);
        bool take_refl = length(refr) < 1e-3f;
        if (!take_refl) take_refl = randf(P.seed) < fr;
        if (take_refl) {
            R.nd = refl;
            R.no = hp + nf * 1e-3f;
            P.skip_sky = false;
            P.prev_lambert = false;
        } else {
            R.nd = normalize(refr);
            R.no = hp - nf * 1e-3f;
        }
        R.alive = true;
    }
    if (b >= 2u && F.rr_enabled) {
        T p = clampf(maxc(P.T_), 0.05f, 1.0f);
        if (randf(P.seed) > p) R.alive = false;
        else P.T_ = P.T_ / p;
    }
}

// One step i of transmittance (PathTrace.slang:5089-5196) given the closest
// hit h of the ray (o, d): returns true when t is final (a miss or a hit
// beyond t_max keeps it, an opaque hit, the step limit or total internal
// reflection zero it); otherwise tints t by (1 - F) * albedo and moves the
// ray through the dielectric surface.
template <class T>
__device__ __forceinline__ bool transmit_step(const SceneView& S, const FrameArgs& F, HitRec h, uint32_t i, V3<T>& o,
                                              V3<T>& d, T& t_max, V3<T>& t) {
    if (h.code == kMiss || h.t >= primal(t_max)) return true;
    const SurfaceT<T> sf = surface_of(S, o, d, h);
    if (sf.mat != 2u || i == F.refract_bounces) {
        t = to<T>(f3{0.0f, 0.0f, 0.0f});
        return true;
    }
    T ndi = dot(sf.n, d);
    V3<T> nf = (ndi < 0.0f) ? sf.n : -sf.n;
    float ia = (ndi < 0.0f) ? 1.0f : sf.ior;
    float ib = (ndi < 0.0f) ? sf.ior : 1.0f;
    float eta = ia / ib;
    T ci = fabs_(dot(nf, -d));
    float q = (ia - ib) / (ia + ib);
    float r0 = q * q;
    T m = 1.0f - ci;
    T fr = r0 + ((((1.0f - r0) * m) * m * m) * m) * m;
    V3<T> rf = refract(d, nf, eta);
    if (length(rf) < 1e-3f) {
        t = to<T>(f3{0.0f, 0.0f, 0.0f});
        return true;
    }
    t = t * (to<T>(sf.albedo) * (1.0f - fr));
    T adv = sf.t + 1e-3f;
    o = o + d * adv;
    d = normalize(rf);
    t_max = fmax_(t_max - adv, T(0.0f));
    return false;
}

// transmittance: closest hits along the ray, refracting through dielectrics,
// at most refract_bounces + 1 queries.
template <bool COUNT>
__device__ __forceinline__ f3 transmittance(const SceneView& S, const FrameArgs& F, f3 o, f3 d, float t_max,
                                            Stack& st, TraceCounters& tc, uint32_t& steps) {
    f3 t{1.0f, 1.0f, 1.0f};
    for (uint32_t i = 0; i <= F.refract_bounces; ++i) {
        steps++;
        HitRec h = trace_closest<COUNT>(S, o, d, st, tc);
        if (transmit_step(S, F, h, i, o, d, t_max, t)) return t;
    }
    return f3{0.0f, 0.0f, 0.0f};
}

// Contribution of a shadow record with transmittance tr (tr > 0 somewhere).
template <class T>
__device__ __forceinline__ V3<T> shadow_value(const FrameArgs& F, int kind, V3<T> pre, T ndl, V3<T> brdf, T pdf,
                                              T w, bool clamp, V3<T> tr) {
    V3<T> c;
    if (kind == 0) c = (((pre * tr) * brdf) * ndl) / pdf;
    else c = ((((pre * tr) * brdf) * ndl) * w) / pdf;
    return clamp ? clamp_indirect(F, c) : c;
}

// Final value of a frame from the sample sum: spp average, NaN/Inf/negative
// sanitising (PathTrace.slang:7722-7734).
__device__ __forceinline__ f3 frame_value(const FrameArgs& F, f3 sum) {
    f3 fr = sum / (float)F.spp;
    float o[3] = {fr.x, fr.y, fr.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = (isnan(o[k]) || isinf(o[k])) ? 0.0f : fmaxf(o[k], 0.0f);
    return f3{o[0], o[1], o[2]};
}

// Running-mean / EMA accumulation (PathTrace.slang:7736-7781).
__device__ __forceinline__ float4 accumulate(const FrameArgs& F, float4 prev, bool reset, f3 v) {
    if (reset) prev = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (isnan(prev.x) || isinf(prev.x) || isnan(prev.y) || isinf(prev.y) || isnan(prev.z) || isinf(prev.z) ||
        isnan(prev.w) || isinf(prev.w))
        prev = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float ema = fminf(fmaxf(F.ema_alpha, 0.0f), 0.999f);
    if (prev.w < 0.5f || ema <= 0.0f) {
        const float nn = prev.w + 1.0f;
        return make_float4((prev.x * prev.w + v.x) / nn, (prev.y * prev.w + v.y) / nn, (prev.z * prev.w + v.z) / nn,
                           nn);
    }
    return make_float4(prev.x * ema + v.x * (1.0f - ema), prev.y * ema + v.y * (1.0f - ema),
                       prev.z * ema + v.z * (1.0f - ema), fminf(prev.w + 1.0f, 1.0f / fmaxf(1.0f - ema, 1e-3f)));
}

}  // namespace dmt
