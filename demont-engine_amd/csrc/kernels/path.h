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
struct ShadowRec {
    f3 o, d;
    float t_max;
    f3 pre;        // throughput * Le (light) or throughput * env colour
    float ndl;
    f3 brdf;
    float pdf;
    float w;       // MIS weight (env)
    bool clamp;    // clampIndirect applies (light NEE at b > 0; env always)
};

struct PathVertexState {
    f3 T, L;
    float prev_pdf;
    bool skip_sky, prev_lambert;
    uint32_t seed;
};

struct VertexResult {
    bool alive;                  // the path scatters (before the depth limit)
    f3 no, nd;                   // next ray
};

// Shades the vertex at bounce b for the ray (o, d) with closest hit h.
// Shadow queries go to sink.emit(kind, record) as soon as they are formed
// (light NEE before env NEE, the order their contributions are summed).
template <class Sink>
__device__ __forceinline__ void shade_vertex(const SceneView& S, const FrameArgs& F, uint32_t b, f3 o, f3 d,
                                             HitRec h, PathVertexState& P, VertexResult& R, Sink& sink) {
    const bool mis_on = F.mis_enabled != 0u;
    R.alive = false;
    R.no = f3{0.0f, 0.0f, 0.0f};
    R.nd = f3{0.0f, 1.0f, 0.0f};
    if (h.code == kMiss) {
        if (mis_on && P.prev_lambert) {
            // transmittance(ro, rd, 1e30) would re-trace this missed ray and
            // return exactly 1 (PathTrace.slang:6081-6084).
            float pe = pdf_env(S, d);
            float w = mis_balance(P.prev_pdf, pe);
            P.L = P.L + clamp_indirect(F, (P.T * sky_color(S, d)) * w);
        } else if (!P.skip_sky) {
            f3 c = P.T * sky_color(S, d);
            P.L = P.L + ((b == 0u) ? c : clamp_indirect(F, c));
        }
        return;
    }
    const Surface sf = surface_at(S, o, d, h);
    const f3 hp = o + d * h.t;
    const float ndi = dot(sf.n, d);
    const f3 nf = (ndi < 0.0f) ? sf.n : -sf.n;
    if (any_pos(sf.emission)) {
        f3 c = P.T * sf.emission;
        P.L = P.L + ((b == 0u) ? c : clamp_indirect(F, c));
    }
    if (sf.mat == 0u) {
        if (S.light_count > 0u) {
            LightSample ls = sample_light(S, F, hp, nf, P.seed);
            float ndl = dot(nf, ls.dir);
            if (ls.pdf > 0.0f && ndl > 0.0f) {
                ShadowRec r;
                r.o = hp + nf * 1e-3f;
                r.d = ls.dir;
                r.t_max = fmaxf(ls.dist - 1e-3f, 0.0f);
                r.pre = P.T * ls.Le;
                r.ndl = ndl;
                r.brdf = sf.albedo / kPi;
                r.pdf = ls.pdf;
                r.w = 1.0f;
                r.clamp = b > 0u;
                sink.emit(0, r, P);
            }
        }
        if (S.env_present) {
            float x0 = randf(P.seed);
            float x1 = randf(P.seed);
            EnvSample e = sample_env(S, x0, x1);
            float ndl = dot(nf, e.dir);
            if (e.pdf > 0.0f && ndl > 0.0f) {
                float pb = ndl * (1.0f / kPi);
                bool fire = mis_on && (b + 1u) < F.max_bounces;
                ShadowRec r;
                r.o = hp + nf * 1e-3f;
                r.d = e.dir;
                r.t_max = 1e30f;
                r.pre = P.T * e.color;
                r.ndl = ndl;
                r.brdf = sf.albedo / kPi;
                r.pdf = e.pdf;
                r.w = fire ? mis_balance(e.pdf, pb) : 1.0f;
                r.clamp = true;
                sink.emit(1, r, P);
            }
            P.skip_sky = !mis_on;
        } else {
            // No env map: ambient albedo * skyColorBase(nf) (PathTrace.slang:6667),
            // added after the light NEE term.
            float tt = clampf(0.5f * (nf.y + 1.0f), 0.0f, 1.0f);
            ShadowRec r;
            r.o = hp;
            r.d = f3{0.0f, 1.0f, 0.0f};
            r.t_max = 0.0f;
            r.pre = (P.T * sf.albedo) * lerp3(f3{0.5f, 0.55f, 0.62f}, f3{1.0f, 1.0f, 1.0f}, tt);
            r.ndl = 0.0f;
            r.brdf = f3{0.0f, 0.0f, 0.0f};
            r.pdf = -1.0f;
            r.w = 0.0f;
            r.clamp = false;
            sink.emit(1, r, P);
            P.skip_sky = true;
        }
        P.T = P.T * sf.albedo;
        R.no = hp + nf * 1e-3f;
        R.nd = cos_sample(nf, P.seed);
        if (mis_on && S.env_present) {
            P.prev_pdf = fmaxf(0.0f, dot(nf, R.nd)) * (1.0f / kPi);
            P.prev_lambert = true;
        } else {
            P.prev_lambert = false;
        }
        R.alive = true;
    } else if (sf.mat == 1u) {
        f3 r = reflect(d, nf);
        f3 cs = cos_sample(nf, P.seed);
        r = normalize(r + (cs - nf) * sf.rough);
        P.T = P.T * sf.albedo;
        R.no = hp + nf * 1e-3f;
        R.nd = r;
        P.skip_sky = false;
        P.prev_lambert = false;
        R.alive = true;
    } else {
        float ia = (ndi < 0.0f) ? 1.0f : sf.ior;
        float ib = (ndi < 0.0f) ? sf.ior : 1.0f;
        float eta = ia / ib;
        float ci = fabsf(dot(nf, -d));
        float r0 = (ia - ib) / (ia + ib);
        r0 = r0 * r0;
        float m = 1.0f - ci;
        float fr = r0 + ((((1.0f - r0) * m) * m * m) * m) * m;
        f3 refl = reflect(d, nf);
        f3 refr = refract(d, nf, eta);
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
    if (b >= 2u) {
        float p = clampf(maxc(P.T), 0.05f, 1.0f);
        if (randf(P.seed) > p) R.alive = false;
        else P.T = P.T / p;
    }
}

// One step i of transmittance (PathTrace.slang:5089-5196) given the closest
// hit h of the ray (o, d): returns true when t is final (a miss or a hit
// beyond t_max keeps it, an opaque hit, the step limit or total internal
// reflection zero it); otherwise tints t by (1 - F) * albedo and moves the
// ray through the dielectric surface.
__device__ __forceinline__ bool transmit_step(const SceneView& S, const FrameArgs& F, HitRec h, uint32_t i, f3& o,
                                              f3& d, float& t_max, f3& t) {
    if (h.code == kMiss || h.t >= t_max) return true;
    const Surface sf = surface_at(S, o, d, h);
    if (sf.mat != 2u || i == F.refract_bounces) {
        t = f3{0.0f, 0.0f, 0.0f};
        return true;
    }
    float ndi = dot(sf.n, d);
    f3 nf = (ndi < 0.0f) ? sf.n : -sf.n;
    float ia = (ndi < 0.0f) ? 1.0f : sf.ior;
    float ib = (ndi < 0.0f) ? sf.ior : 1.0f;
    float eta = ia / ib;
    float ci = fabsf(dot(nf, -d));
    float q = (ia - ib) / (ia + ib);
    float r0 = q * q;
    float m = 1.0f - ci;
    float fr = r0 + ((((1.0f - r0) * m) * m * m) * m) * m;
    f3 rf = refract(d, nf, eta);
    if (length(rf) < 1e-3f) {
        t = f3{0.0f, 0.0f, 0.0f};
        return true;
    }
    t = t * (sf.albedo * (1.0f - fr));
    float adv = h.t + 1e-3f;
    o = o + d * adv;
    d = normalize(rf);
    t_max = fmaxf(t_max - adv, 0.0f);
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
__device__ __forceinline__ f3 shadow_value(const FrameArgs& F, int kind, f3 pre, float ndl, f3 brdf, float pdf,
                                           float w, bool clamp, f3 tr) {
    f3 c;
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
