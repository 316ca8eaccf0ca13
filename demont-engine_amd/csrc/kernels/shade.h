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

constexpr uint32_t kMaxHdriLights = 8;   // Engine::kMaxHdriLights (src/engine/Engine.h:1507)

// Sky lights of a frame: the sun and moon of the gradient sky (no env map;
// sun_and_mode.xyz, sun_extra, sun_extra2.x, moon_dir_phase, moon_extra) and
// the directional lights extracted from an env map's bright clusters
// (hdri_lights_dir / hdri_lights_col / hdri_lights_count).  All zero: none.
struct SkyArgs {
    f3 sun;                                // unit direction to the sun
    float sun_size, sun_dist, sun_vflat;   // r_sun_size, distance ratio, vertical flattening (1: round)
    float sun_bright;                      // r_sun_disc_brightness
    f3 moon;                               // unit direction to the moon
    float moon_phase;                      // phase angle (rad)
    float moon_size, moon_dist, moon_bright;
    uint32_t hdri_count;                   // <= kMaxHdriLights
    float4 hdri_dir[kMaxHdriLights];       // xyz direction, w pick probability
    float4 hdri_col[kMaxHdriLights];       // rgb flux
};

// Extensions a frame-completion kernel variant carries: the frame's sky
// lights (four shadow-record kinds), the scene's SDF clusters, its PBR
// material maps and its water.
constexpr uint32_t kExtSky = 1u, kExtSdf = 2u, kExtPbr = 4u, kExtWater = 8u;
constexpr uint32_t kExtVariants = 16u;

// Camera-sample features and sky lights of a frame (all off: the plain
// camera ray and no sky lights).
struct FeatureArgs {
    float aperture, focal, rot, cat_eye;   // thin lens (dof_params, dof_bokeh_params.x)
    int blades;                            // < 3: round aperture
    uint32_t motion_blur;                  // draw a shutter time per sample
    uint32_t spectral;                     // draw a hero wavelength per sample
    uint32_t fixed_jitter;                 // sample 0 takes jit_x, jit_y (halton_jitter) instead of two draws
    float jit_x, jit_y;
    uint32_t ext;                          // kExtSky: sky lights (SkyArgs below); kExtSdf: the
                                           // scene's SDF clusters -- the kernel variant with their code
    uint32_t cam_in_water;                 // camera paths start in the water medium (water_state.x)
    SkyArgs sky_args;
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

// sampleSunDisc (PathTraceMath.slang:55-67): a direction uniform over the
// disc of angular radius 0.00465 rad around dir (two draws).
__device__ __forceinline__ f3 sample_sun_disc(f3 dir, uint32_t& seed) {
    const float u1 = randf(seed);
    const float u2 = randf(seed);
    const float r = sqrtf(u1) * 0.00465f;
    const float phi = 6.28318530718f * u2;
    float sp, cp;
    cephes_sincos(phi, &sp, &cp);
    const float dx = cp * r, dy = sp * r;
    const f3 ax = fabsf(dir.y) < 0.99f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
    const f3 t = normalize(cross(ax, dir));
    const f3 b = cross(dir, t);
    return normalize((dir + t * dx) + b * dy);
}

// The sun's NEE radiance at elevation sin(elev) = dir.y: colour warmed and
// dimmed by the air mass 1 / max(elev, 0.04) (PathTrace.slang:6577-6600).
__device__ __forceinline__ f3 sun_radiance(float elev) {
    const float am = 1.0f / fmaxf(elev, 0.04f);
    const float trans_atm = c_exp(-0.30f * am);
    const float warmth = saturate((am - 1.0f) * 0.18f);
    const f3 col = lerp3(f3{1.00f, 0.97f, 0.92f}, f3{1.55f, 0.55f, 0.20f}, warmth);
    return col * (80.0f * trans_atm);
}

// The moon's NEE radiance (PathTrace.slang:6636-6654), phase_b the lit
// fraction 0.5 (1 - cos(phase)).
__device__ __forceinline__ f3 moon_radiance(float elev, float phase_b) {
    const float m_am = 1.0f / fmaxf(elev, 0.04f);
    const float m_atm = c_exp(-0.30f * m_am);
    const float m_warmth = saturate((m_am - 1.0f) * 0.18f);
    const f3 m_col = lerp3(f3{0.92f, 0.94f, 1.00f}, f3{1.20f, 0.65f, 0.40f}, m_warmth);
    return m_col * ((0.3f * phase_b) * m_atm);
}

// sunDisc (PathTrace.slang:3112-3192): the 0.55-degree disc (an oval when
// sun_vflat < 0.999) and its halo, warmed and dimmed with the air mass.
__device__ __forceinline__ f3 sun_disc(const SkyArgs& K, f3 rd) {
    const f3 sd = K.sun;
    if (sd.y < -0.10f) return f3{0.0f, 0.0f, 0.0f};
    const float ang = dot(rd, sd);
    const float sun_mult = fmaxf(K.sun_size * K.sun_dist, 0.1f);
    const float r_deg = 0.55f * sun_mult;
    const float cos_r = c_cos(r_deg * 0.0174533f);
    const float hz = saturate(1.0f - sd.y);
    const float halo_deg = 2.0f + hz * (8.0f - 2.0f);
    const float cos_halo = c_cos(halo_deg * 0.0174533f);
    if (ang <= cos_halo) return f3{0.0f, 0.0f, 0.0f};
    const float vflat = K.sun_vflat;
    bool inside = false;
    if (ang > cos_r) {
        if (!(vflat < 0.999f)) {
            inside = true;
        } else {
            const f3 lr = normalize(cross(f3{0.0f, 1.0f, 0.0f}, sd));
            const f3 lu = cross(sd, lr);
            const f3 delta = rd - sd * ang;
            const float h = dot(delta, lr), v = dot(delta, lu);
            const float rh = r_deg * 0.0174533f;
            const float rv = rh * vflat;
            inside = (h * h) / (rh * rh) + (v * v) / (rv * rv) <= 1.0f;
        }
    }
    const float am = 1.0f / fmaxf(sd.y, 0.04f);
    const float atm = c_exp(-0.30f * am);
    const float warmth = saturate((am - 1.0f) * 0.18f);
    const f3 col = lerp3(f3{1.00f, 0.98f, 0.94f}, f3{1.55f, 0.55f, 0.20f}, warmth);
    const float t_h = (ang - cos_halo) / (1.0f - cos_halo);
    const float falloff = 8.0f + hz * (3.0f - 8.0f);
    const float halo = c_pow(t_h, falloff);
    const float bright = fmaxf(K.sun_bright, 0.0f);
    const f3 halo_col = ((col * halo) * atm) * 4.0f;
    if (inside) return (col * (250.0f * atm) + halo_col) * bright;
    return halo_col * bright;
}

__device__ __forceinline__ f3 moon_texel(const SceneView& S, int x, int y) {
    return S.moon ? xyz(S.moon[(size_t)y * S.moon_w + x]) : f3{0.0f, 0.0f, 0.0f};
}

// moonDisc (PathTrace.slang:3212-3353): disc and halo at 0.55 / 1.5 degrees
// times the size, the textured surface lit across the terminator by the sun,
// limb darkening, the lunar eclipse near the anti-solar point.  Without a
// moon texture the surface reads 0.
__device__ __forceinline__ f3 moon_disc(const SceneView& S, const SkyArgs& K, f3 rd) {
    const f3 md = K.moon;
    if (md.y < -0.10f) return f3{0.0f, 0.0f, 0.0f};
    const float ang = dot(rd, md);
    const float size_mult = fmaxf(K.moon_size * K.moon_dist, 0.1f);
    const float r_deg = 0.55f * size_mult;
    const float halo_deg = 1.5f * size_mult;
    const float cos_r = c_cos(r_deg * 0.0174533f);
    const float cos_halo = c_cos(halo_deg * 0.0174533f);
    if (ang <= cos_halo) return f3{0.0f, 0.0f, 0.0f};
    const f3 tint{0.92f, 0.94f, 1.00f};
    const float m_am = 1.0f / fmaxf(md.y, 0.04f);
    const float m_atm = c_exp(-0.30f * m_am);
    const float warmth = saturate((m_am - 1.0f) * 0.18f);
    const f3 col_atm = lerp3(f3{1.0f, 1.0f, 1.0f}, f3{1.20f, 0.65f, 0.40f}, warmth);
    const float t_h = (ang - cos_halo) / (1.0f - cos_halo);
    const float halo = c_pow(t_h, 4.0f) * 0.6f;
    const f3 tc = tint * col_atm;
    if (ang > cos_r) {
        const float t_d = (ang - cos_r) / (1.0f - cos_r);
        const f3 ax = fabsf(md.y) < 0.99f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
        const f3 ta = normalize(cross(ax, md));
        const f3 tb = cross(md, ta);
        const float disc_scale = 1.0f / fmaxf(c_sin(r_deg * 0.0174533f), 1e-5f);
        const float du = dot(rd, ta) * disc_scale;
        const float dv = dot(rd, tb) * disc_scale;
        const float r2 = du * du + dv * dv;
        const float zz = sqrtf(fmaxf(1.0f - r2, 0.0f));
        const f3 n_world = (ta * du + tb * dv) - md * zz;
        const float lit = saturate(dot(n_world, K.sun) * 10.0f + 0.05f);
        if (lit <= 0.0f) return ((tc * halo) * m_atm) * 0.05f;
        const float anti_sun_ang = c_acos(saturate(-dot(md, K.sun)));
        const float eclipse = 1.0f - smoothstep(0.0122f, 0.0218f, anti_sun_ang);
        const float lon = c_atan2(du, zz);
        const float lat = c_asin(fminf(fmaxf(dv, -1.0f), 1.0f));
        const float u_tex = saturate(0.5f + lon / 6.28318530718f);
        const float v_tex = saturate(0.5f - lat / 3.14159265358979f);
        const int W = (int)S.moon_w, H = (int)S.moon_h;
        f3 surface{0.0f, 0.0f, 0.0f};
        if (S.moon) {
            const float fx = u_tex * (float)W - 0.5f, fy = v_tex * (float)H - 0.5f;
            const int ix = (int)floorf(fx), iy = (int)floorf(fy);
            const float ax2 = fx - (float)ix, ay2 = fy - (float)iy;
            const int x0 = min(max(ix, 0), W - 1), x1 = min(max(ix + 1, 0), W - 1);
            const int y0 = min(max(iy, 0), H - 1), y1 = min(max(iy + 1, 0), H - 1);
            surface = lerp3(lerp3(moon_texel(S, x0, y0), moon_texel(S, x1, y0), ax2),
                            lerp3(moon_texel(S, x0, y1), moon_texel(S, x1, y1), ax2), ay2);
        }
        const float limb = 0.55f + 0.45f * sqrtf(saturate(t_d));
        const float bright_mult = fmaxf(K.moon_bright, 0.0f);
        const float disc_b = 3.5f * bright_mult;
        const f3 lit_col = (tc * surface) * (((disc_b * lit) * limb) * m_atm) +
                           (((tc * halo) * lit) * m_atm) * (0.6f * bright_mult);
        const f3 eclipse_col = (f3{0.55f, 0.15f, 0.04f} * surface) * ((limb * m_atm) * 1.2f);
        return lerp3(lit_col, eclipse_col, eclipse);
    }
    const float phase_b = 0.5f * (1.0f - c_cos(K.moon_phase));
    return (((tc * halo) * phase_b) * m_atm) * 0.3f;
}

// The gradient sky with the sun and moon discs (skyColor without an env map,
// PathTrace.slang:3451-3452: base + moonDisc + sunDisc).
__device__ __forceinline__ f3 gradient_sky(const SceneView& S, const SkyArgs& K, f3 rd) {
    const float t = saturate(0.5f * (rd.y + 1.0f));
    const f3 base = lerp3(f3{0.5f, 0.55f, 0.62f}, f3{1.0f, 1.0f, 1.0f}, t);
    return (base + moon_disc(S, K, rd)) + sun_disc(K, rd);
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

// isExtractedClusterDir (PathTrace.slang:3679-3696): the direction's texel
// is one the env map's importance CDF leaves out (the conditional CDF does not
// rise there): an extracted cluster, lit by its own directional NEE.
__device__ __forceinline__ bool is_cluster_dir(const SceneView& S, const SkyArgs& K, f3 d) {
    if (!S.env_present || K.hdri_count == 0u) return false;
    const float th = c_acos(clampf(d.y, -1.0f, 1.0f));
    const float st = c_sin(th);
    if (st < 1e-6f) return false;
    const float ph = c_atan2(d.z, d.x);
    const uint32_t W = S.env_w, H = S.env_h;
    const float u = (ph + kPi) * (1.0f / 6.28318530717958f);
    const float v = th * (1.0f / kPi);
    const uint32_t ui = min((uint32_t)(u * (float)W), W - 1u);
    const uint32_t vi = min((uint32_t)(v * (float)H), H - 1u);
    const float* row = S.env_cond + (size_t)vi * W;
    const float c_prev = ui > 0u ? row[ui - 1u] : 0.0f;
    return row[ui] <= c_prev;
}

struct HdriSample {
    f3 dir, color;
    float pdf;
};

// sampleHdriLight (PathTrace.slang:3749-3768): one cluster picked by its
// probability (w of hdri_dir) with one draw; the last when rounding leaves
// the running sum below the draw.
__device__ __forceinline__ HdriSample sample_hdri(const SceneView& S, const SkyArgs& K, uint32_t& seed) {
    HdriSample s{f3{0.0f, 1.0f, 0.0f}, f3{0.0f, 0.0f, 0.0f}, 0.0f};
    if (K.hdri_count == 0u) return s;
    const float u = randf(seed);
    float acc = 0.0f;
    bool found = false;
    float4 dsel = K.hdri_dir[0], csel = K.hdri_col[0];
    // unrolled, every index constant: the arrays stay out of private memory
#pragma unroll
    for (uint32_t i = 0; i < kMaxHdriLights; ++i) {
        if (!found && i < K.hdri_count) {
            acc = acc + K.hdri_dir[i].w;
            dsel = K.hdri_dir[i];
            csel = K.hdri_col[i];
            found = u <= acc;
        }
    }
    s.dir = xyz(dsel);
    s.color = xyz(csel) * S.env_intensity;
    s.pdf = fmaxf(dsel.w, 1e-6f);
    return s;
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
