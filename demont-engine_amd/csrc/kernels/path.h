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
#include "pbr.h"
#include "water.h"

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

// cauchyIOR (PathTraceMath.slang:111-119): n = B + C / lambda^2 through
// (base, 587.6 nm), lambda in micrometres.
__device__ __forceinline__ float cauchy_ior(float base, float cc, float lambda_nm) {
    const float lum = lambda_nm * 0.001f;
    const float l2 = lum * lum;
    const float ld2 = 0.5876f * 0.5876f;
    const float B = base - cc / ld2;
    return B + cc / l2;
}

// wymanGauss, cieXYZFit, xyzToLinearSRGB, spectralWeightSRGB
// (PathTraceMath.slang:140-192): the Wyman-Sloan-Shirley fit of the CIE 1931
// matching functions, to linear sRGB, over the band means.
__device__ __forceinline__ float wyman(float x, float mu, float s1, float s2) {
    const float t = (x - mu) * ((x < mu) ? s1 : s2);
    return c_exp((-0.5f * t) * t);
}
__device__ __forceinline__ f3 spectral_weight(float l) {
    const float X = (1.056f * wyman(l, 599.8f, 0.0264f, 0.0323f) + 0.362f * wyman(l, 442.0f, 0.0624f, 0.0374f)) -
                    0.065f * wyman(l, 501.1f, 0.0490f, 0.0382f);
    const float Y = 0.821f * wyman(l, 568.8f, 0.0213f, 0.0247f) + 0.286f * wyman(l, 530.9f, 0.0613f, 0.0322f);
    const float Z = 1.217f * wyman(l, 437.0f, 0.0845f, 0.0278f) + 0.681f * wyman(l, 459.0f, 0.0385f, 0.0725f);
    const float r = (3.2404542f * X - 1.5371385f * Y) - 0.4985314f * Z;
    const float g = (-0.9692660f * X + 1.8760108f * Y) + 0.0415560f * Z;
    const float b = (0.0556434f * X - 0.2040259f * Y) + 1.0572252f * Z;
    return f3{r * 2.730792f, g * 3.443936f, b * 3.606855f};
}

// ptConcentricDisk (PathTraceMath.slang:223-236).
__device__ __forceinline__ void concentric_disk(float u1, float u2, float& x, float& y) {
    const float a = 2.0f * u1 - 1.0f;
    const float b = 2.0f * u2 - 1.0f;
    const float qpi = 3.14159265358979f / 4.0f, hpi = 3.14159265358979f / 2.0f;
    float r, phi;
    if (fabsf(a) > fabsf(b)) {
        r = a;
        phi = qpi * (b / fmaxf(fabsf(a), 1e-6f));
    } else {
        r = b;
        phi = hpi - qpi * (a / fmaxf(fabsf(b), 1e-6f));
    }
    float sn, cs;
    cephes_sincos(phi, &sn, &cs);
    x = r * cs;
    y = r * sn;
}

// ptPolygonSample (PathTraceMath.slang:268-286): a wedge of the regular
// N-gon picked by u1, then the sqrt-barycentric point of that triangle.
__device__ __forceinline__ void polygon_sample(float u1, float u2, int blades, float rot, float& x, float& y) {
    const float n = (float)blades;
    const float wedge = 6.28318530717958f / n;
    const float kf = floorf(u1 * n);
    const float ku = u1 * n - kf;
    const float a0 = rot + kf * wedge;
    const float a1 = a0 + wedge;
    float s0, c0, s1, c1;
    cephes_sincos(a0, &s0, &c0);
    cephes_sincos(a1, &s1, &c1);
    const float su = sqrtf(ku);
    const float w1 = 1.0f - u2;
    x = su * (w1 * c0 + u2 * c1);
    y = su * (w1 * s0 + u2 * s1);
}

// ptCatEye (PathTraceMath.slang:296-310): the radial component of the lens
// sample narrowed with the image eccentricity |uv|.
__device__ __forceinline__ void cat_eye(float& x, float& y, float ux, float uy, float strength) {
    const float len = sqrtf(ux * ux + uy * uy);
    const float ecc = fminf(fmaxf(len, 0.0f), 1.0f);
    const float squeeze = strength * ecc;
    if (squeeze < 1e-5f) return;
    float rx = 1.0f, ry = 0.0f;
    if (ecc > 1e-5f) { rx = ux / len; ry = uy / len; }
    const float tx = -ry, ty = rx;
    float ar = x * rx + y * ry;
    const float at = x * tx + y * ty;
    ar = ar * (1.0f - 0.7f * squeeze);
    x = rx * ar + tx * at;
    y = ry * ar + ty * at;
}

// Camera ray of sample s of pixel (x, y) with the camera-sample features
// (PathTrace.slang:5729-5816): the jitter, then, each only when its feature
// is on, the hero wavelength (stratified over the spp samples), the shutter
// time and the lens point.  Without them: camera_ray, time 1, lambda -1.
// With fixed_jitter (the denoiser's Halton offset, :5730-5731) sample 0
// takes (jit_x, jit_y) and draws nothing for it.
__device__ __forceinline__ void camera_sample(const FrameArgs& F, const FeatureArgs& X, uint32_t x, uint32_t y,
                                              uint32_t s, uint32_t& seed, f3& o, f3& d, float& time, float& lambda) {
    float jx, jy;
    if (X.fixed_jitter && s == 0u) {
        jx = X.jit_x;
        jy = X.jit_y;
    } else {
        jx = randf(seed);
        jy = randf(seed);
        jx = jx - 0.5f;
        jy = jy - 0.5f;
    }
    const float ux = ((((float)x + 0.5f) + jx) / (float)F.width) * 2.0f - 1.0f;
    float uy = ((((float)y + 0.5f) + jy) / (float)F.height) * 2.0f - 1.0f;
    uy = -uy;
    lambda = -1.0f;
    if (X.spectral) {
        const float u = ((float)s + randf(seed)) / (float)F.spp;
        lambda = 380.0f + u * (730.0f - 380.0f);
    }
    time = 1.0f;
    if (X.motion_blur) time = randf(seed);
    o = F.pos;
    d = normalize((F.fwd + F.right * ((ux * F.aspect) * F.fovtan)) + F.up * (uy * F.fovtan));
    if (X.aperture > 0.0f) {
        const float u1 = randf(seed);
        const float u2 = randf(seed);
        float ax, ay;
        if (X.blades >= 3) polygon_sample(u1, u2, X.blades, X.rot, ax, ay);
        else concentric_disk(u1, u2, ax, ay);
        if (X.cat_eye > 0.0f) cat_eye(ax, ay, ux, uy, X.cat_eye);
        const f3 lens = F.right * (ax * X.aperture) + F.up * (ay * X.aperture);
        const f3 fp = o + d * X.focal;
        o = o + lens;
        d = normalize(fp - o);
    }
}

// A shadow query and the factors of its contribution.  kind 0 = light NEE;
// 1 = env NEE, or without an env map the ambient term (pdf < 0, added
// without tracing); 2 = the HDRI cluster NEE (env map) or the sun NEE (no env
// map); 3 = the moon NEE.  A vertex's terms are summed in the order
// 0, 1, 2 with an env map and 0, 2, 3, 1 without one (shade_kind_order).
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
    float time = 1.0f;      // shutter time of the sample (1: no motion blur)
    float lambda = -1.0f;   // hero wavelength in nm (< 0: RGB)
    float cone = 0.0f;      // ray-cone width at the current origin (PBR frames)
    bool in_water = false;  // the ray travels in the water medium (water frames)
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
    // The same vertex with the ray (o, d) that reached it, its closest hit h
    // and the path state P it arrived with (the adjoint records its metal
    // vertices here, to replay them in dual numbers from this state).
    template <class T>
    __device__ __forceinline__ void on_vertex(uint32_t b, const SurfaceT<T>& sf, V3<T> o, V3<T> d, HitRec h,
                                              const PathVertexStateT<T>& P) {}
};

// The k-th kind whose shadow record a vertex's terms are summed in
// (k = 0 .. 3; kinds a vertex did not emit are skipped).
__device__ __forceinline__ uint32_t shade_kind_order(bool env, uint32_t k) {
    return env ? k : ((0x1320u >> (4u * k)) & 15u);   // no env map: 0, 2, 3, 1
}

// The sky of a missed ray: skyColor, with the sun and moon discs when SKY.
template <bool SKY, class T>
__device__ __forceinline__ V3<T> miss_sky(const SceneView& S, const SkyArgs* K, V3<T> d) {
    if constexpr (SKY) {
        if (!S.env_present) return gradient_sky(S, *K, d);
    }
    return sky_color(S, d);
}

// Shades the vertex at bounce b for the ray (o, d) with closest hit h.
// Shadow queries go to sink.emit(kind, record) as soon as they are formed,
// in the order their contributions are summed.  SKY (K non-null): the
// frame's sky lights -- the sun and moon NEE of the no-env Lambert branch
// with their discs in the sky (PathTrace.slang:6546-6667), the HDRI cluster
// NEE of the env branch and the cluster test of the MIS miss (:6493-6518,
// :6074); SKY frames are float only.  SDF: h may be an SDF cluster's hit
// (the feature path's traversal).
template <bool SKY = false, bool SDF = false, bool PBR = false, bool WATER = false, class T, class Sink>
__device__ __forceinline__ void shade_vertex(const SceneView& S, const FrameArgs& F, uint32_t b, V3<T> o, V3<T> d,
                                             HitRec h, PathVertexStateT<T>& P, VertexResultT<T>& R, Sink& sink,
                                             const SkyArgs* K = nullptr) {
    const bool mis_on = F.mis_enabled != 0u;
    R.alive = false;
    R.no = to<T>(f3{0.0f, 0.0f, 0.0f});
    R.nd = to<T>(f3{0.0f, 1.0f, 0.0f});
    if constexpr (WATER) {
        static_assert(sizeof(T) == sizeof(float), "water frames are float only");
        if (P.in_water) {
            if (h.code != kMiss) {
                // Beer's law over the segment just traced (PathTrace.slang:6028-6031)
                P.T_ = P.T_ * water_absorb(S, h.t);
            } else {
                // escaped under water: the deep-water tint (:6046-6052)
                const f3 v = P.T_ * S.deep_color;
                P.L = P.L + ((b == 0u) ? v : clamp_indirect(F, v));
                return;
            }
        }
    }
    if (h.code == kMiss) {
        if (mis_on && P.prev_lambert) {
            // cluster directions are left to the clusters' own NEE (:6074)
            bool cluster = false;
            if constexpr (SKY) cluster = is_cluster_dir(S, *K, d);
            if (!cluster) {
                // transmittance(ro, rd, 1e30) would re-trace this missed ray and
                // return exactly 1 (PathTrace.slang:6081-6084).
                T pe = pdf_env_of(S, d);
                T w = mis_balance(P.prev_pdf, pe);
                const V3<T> v = (P.T_ * sky_color(S, d)) * w;
                const V3<T> c = clamp_indirect(F, v);
                P.L = P.L + c;
                sink.on_radiance(v, c, true);
            }
        } else if (!P.skip_sky) {
            V3<T> v = P.T_ * miss_sky<SKY>(S, K, d);
            const V3<T> c = (b == 0u) ? v : clamp_indirect(F, v);
            P.L = P.L + c;
            sink.on_radiance(v, c, b != 0u);
        }
        return;
    }
    SurfaceT<T> sf = surface_of<SDF>(S, o, d, h, P.time);
    if constexpr (PBR) {
        // the ray cone grows by this segment, then the hit's material maps
        // apply before the shading normal and the BSDF are formed
        // (PathTrace.slang:6162-6171); PBR frames are float only
        static_assert(sizeof(T) == sizeof(float), "PBR frames are float only");
        P.cone = P.cone + S.cone_spread * sf.t;
        apply_pbr(S, o, d, o + d * sf.t, h, sf, P.time, P.cone);
    }
    sink.on_surface(b, sf);
    sink.on_vertex(b, sf, o, d, h, P);
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
            if constexpr (SKY) {
                if (K->hdri_count > 0u) {
                    // one cluster per sample, contribution / its probability;
                    // unclamped at the primary hit (:6503-6518)
                    const HdriSample hl = sample_hdri(S, *K, P.seed);
                    const float ndl = dot(nf, hl.dir);
                    if (ndl > 0.0f && hl.pdf > 0.0f) {
                        ShadowRecT<T> r;
                        r.o = hp + nf * 1e-3f;
                        r.d = hl.dir;
                        r.t_max = 1e30f;
                        r.pre = P.T_ * hl.color;
                        r.ndl = ndl;
                        r.brdf = sf.albedo / kPi;
                        r.pdf = hl.pdf;
                        r.w = 1.0f;
                        r.clamp = b > 0u;
                        sink.emit(2, r, P);
                    }
                }
            }
            P.skip_sky = !mis_on;
        } else {
            if constexpr (SKY) {
                // the sun's NEE: a shadow ray into its disc, radiance by the
                // air mass, unclamped (:6552-6620)
                const f3 sd = K->sun;
                if (sd.y > 0.0f) {
                    const float ndl = fmaxf(0.0f, dot(nf, sd));
                    const f3 jit = sample_sun_disc(sd, P.seed);
                    if (ndl > 0.0f) {
                        ShadowRecT<T> r;
                        r.o = hp + nf * 1e-3f;
                        r.d = jit;
                        r.t_max = 1e30f;
                        r.pre = P.T_ * sun_radiance(sd.y);
                        r.ndl = ndl;
                        r.brdf = sf.albedo / kPi;
                        r.pdf = 1.0f;
                        r.w = 1.0f;
                        r.clamp = false;
                        sink.emit(2, r, P);
                    }
                }
                // the moon's NEE when it is up and more than 1 % lit (:6624-6660)
                const f3 md = K->moon;
                if (md.y > 0.0f) {
                    const float phase_b = 0.5f * (1.0f - c_cos(K->moon_phase));
                    if (phase_b > 0.01f) {
                        const f3 jit = sample_sun_disc(md, P.seed);
                        const float ndl = fmaxf(0.0f, dot(nf, jit));
                        if (ndl > 0.0f) {
                            ShadowRecT<T> r;
                            r.o = hp + nf * 1e-3f;
                            r.d = jit;
                            r.t_max = 1e30f;
                            r.pre = P.T_ * moon_radiance(md.y, phase_b);
                            r.ndl = ndl;
                            r.brdf = sf.albedo / kPi;
                            r.pdf = 1.0f;
                            r.w = 1.0f;
                            r.clamp = false;
                            sink.emit(3, r, P);
                        }
                    }
                }
            }
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
    } else if (WATER && sf.mat == kMatWater) {
        if constexpr (WATER) {
            // the water surface (:6768-6906 without the FFT ocean and its
            // foam): the wave normal, Schlick with F0 0.02; reflection (or
            // total internal reflection) keeps the medium, refraction
            // crosses it; no NEE, the throughput unchanged
            const float wior = water_ior_of(S, sf.ior);
            const f3 sn = water_normal(S, hp, sf.n);
            const float ndi_s = dot(sn, d);
            const f3 nf_s = (ndi_s < 0.0f) ? sn : -sn;
            const float ia = (ndi_s < 0.0f) ? 1.0f : wior;
            const float ib = (ndi_s < 0.0f) ? wior : 1.0f;
            const float eta = ia / ib;
            const float ci = fabsf(dot(nf_s, -d));
            const float r0 = 0.02f;
            const float m = 1.0f - ci;
            const float fr = r0 + ((((1.0f - r0) * m) * m * m) * m) * m;
            const f3 refl = reflect(d, nf_s);
            const f3 refr = refract(d, nf_s, eta);
            bool take_refl = length(refr) < 1e-3f;
            if (!take_refl) take_refl = randf(P.seed) < fr;
            if (take_refl) {
                R.nd = refl;
                R.no = hp + nf_s * 1e-3f;
                P.skip_sky = false;
                P.prev_lambert = false;
            } else {
                P.in_water = ndi_s < 0.0f;
                R.nd = normalize(refr);
                R.no = hp - nf_s * 1e-3f;
            }
            R.alive = true;
        }
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
        // the hero wavelength's own index (PathTrace.slang:6721-6724)
        const float ior = (P.lambda >= 0.0f) ? cauchy_ior(sf.ior, F.cauchy_c, P.lambda) : sf.ior;
        float ia = (ndi < 0.0f) ? 1.0f : ior;
        float ib = (ndi < 0.0f) ? ior : 1.0f;
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
// ray through the dielectric surface.  SDF as in shade_vertex.  WATER (water
// frames): in_water is the query's medium -- the segment in the water is
// absorbed first, and a water surface transmits (1 - F) with F0 0.02 and
// flips the medium.
template <bool SDF = false, bool WATER = false, class T>
__device__ __forceinline__ bool transmit_step(const SceneView& S, const FrameArgs& F, HitRec h, uint32_t i, V3<T>& o,
                                              V3<T>& d, T& t_max, V3<T>& t, float time, bool& in_water) {
    if constexpr (WATER) {
        // Beer's law along the part of this segment in the water (:5121-5124)
        if (in_water) {
            const float seg = (h.code != kMiss) ? fminf(h.t, primal(t_max)) : primal(t_max);
            t = t * to<T>(water_absorb(S, fmaxf(seg, 0.0f)));
        }
    }
    if (h.code == kMiss || h.t >= primal(t_max)) return true;
    const SurfaceT<T> sf = surface_of<SDF>(S, o, d, h, time);
    const bool water = WATER && sf.mat == kMatWater;
    if ((sf.mat != 2u && !water) || i == F.refract_bounces) {
        t = to<T>(f3{0.0f, 0.0f, 0.0f});
        return true;
    }
    const float ior = water ? water_ior_of(S, sf.ior) : sf.ior;
    T ndi = dot(sf.n, d);
    V3<T> nf = (ndi < 0.0f) ? sf.n : -sf.n;
    float ia = (ndi < 0.0f) ? 1.0f : ior;
    float ib = (ndi < 0.0f) ? ior : 1.0f;
    float eta = ia / ib;
    T ci = fabs_(dot(nf, -d));
    float q = (ia - ib) / (ia + ib);
    float r0 = water ? 0.02f : q * q;
    T m = 1.0f - ci;
    T fr = r0 + ((((1.0f - r0) * m) * m * m) * m) * m;
    V3<T> rf = refract(d, nf, eta);
    if (length(rf) < 1e-3f) {
        t = to<T>(f3{0.0f, 0.0f, 0.0f});
        return true;
    }
    if (water) {
        // water: the Fresnel-transmitted fraction, and the medium flips
        t = t * (1.0f - fr);
        in_water = !in_water;
    } else {
        t = t * (to<T>(sf.albedo) * (1.0f - fr));
    }
    T adv = sf.t + 1e-3f;
    o = o + d * adv;
    d = normalize(rf);
    t_max = fmax_(t_max - adv, T(0.0f));
    return false;
}

template <bool SDF = false, class T>
__device__ __forceinline__ bool transmit_step(const SceneView& S, const FrameArgs& F, HitRec h, uint32_t i, V3<T>& o,
                                              V3<T>& d, T& t_max, V3<T>& t, float time = 1.0f) {
    bool dry = false;
    return transmit_step<SDF, false>(S, F, h, i, o, d, t_max, t, time, dry);
}

// transmittance: closest hits along the ray, refracting through dielectrics,
// at most refract_bounces + 1 queries (at shutter time `time` with MOTION).
template <bool COUNT, bool MOTION = false>
__device__ __forceinline__ f3 transmittance(const SceneView& S, const FrameArgs& F, f3 o, f3 d, float t_max,
                                            Stack& st, TraceCounters& tc, uint32_t& steps, float time = 1.0f) {
    f3 t{1.0f, 1.0f, 1.0f};
    for (uint32_t i = 0; i <= F.refract_bounces; ++i) {
        steps++;
        HitRec h = trace_closest<COUNT, MOTION>(S, o, d, st, tc, time);
        if (transmit_step(S, F, h, i, o, d, t_max, t, time)) return t;
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

// Final value of a frame from the sample sum: spp average, NaN/Inf zeroed,
// negative channels zeroed unless the frame is spectral
// (PathTrace.slang:7722-7734): the hero wavelength's sRGB weights are
// legitimately negative, and those signed values must reach the running mean
// for it to stay unbiased (the display clamps them).
__device__ __forceinline__ f3 frame_value(const FrameArgs& F, f3 sum, bool spectral = false) {
    f3 fr = sum / (float)F.spp;
    float o[3] = {fr.x, fr.y, fr.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = (isnan(o[k]) || isinf(o[k])) ? 0.0f : (spectral ? o[k] : fmaxf(o[k], 0.0f));
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
