// Wavefront kernels of the gfx950 path tracer.
//
//   k_raygen      camera rays (PathTrace.slang:5729-5778), path-state reset
//   k_extend      closest hit for the live ray queue (traceScene, :2331-2693)
//   k_shade       miss / emission / Lambert / metal / dielectric / RR
//                 (:6034-6913); emits shadow queries and the next ray queue
//   k_shadow<0>   light-NEE transmittance (:5089-5196) and contribution (:6341-6366)
//   k_shadow<1>   env-NEE transmittance and MIS-weighted contribution (:6457-6492)
//   k_sample_end  spp average, sanitize, running-mean accumulate (:7707-7781)
//
// Queue appends are wave-aggregated: one atomic per wave64 for each queue,
// slots from a ballot prefix.  Queue order is not deterministic, the image
// is: every path's arithmetic is done by one lane in a fixed order.
#include "wavefront.h"

namespace dmt {

__device__ __forceinline__ uint32_t wave_append(uint32_t* ctr, bool pred) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return 0u;
    const uint32_t lane = __lane_id();
    const uint32_t below = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    const int leader = __ffsll(mask) - 1;
    uint32_t base = 0u;
    if ((int)lane == leader) base = atomicAdd(ctr, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + below;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void add_stat(unsigned long long* dst, uint32_t v) {
    uint32_t s = wave_sum(v);
    if (__lane_id() == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

__device__ __forceinline__ Stack make_stack(uint32_t* lds, uint32_t* spill, uint32_t spill_depth) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    return Stack{lds + threadIdx.x, kBlock, kLdsDepth, spill + (size_t)gid * spill_depth, spill_depth};
}

// ---------------------------------------------------------------------------

// Zeroes the queue counters of a sample; the bounce-0 queue holds every pixel.
__global__ void k_reset_counters(uint32_t* ctr, uint32_t npix) {
    for (uint32_t k = threadIdx.x; k < kCounterSlots; k += blockDim.x) ctr[k] = (k == ctr_rays(0)) ? npix : 0u;
}

__global__ void __launch_bounds__(kBlock) k_raygen(FrameArgs F, PathState ps, RayQueue q, uint32_t s,
                                                   uint32_t* ctr) {
    const uint32_t npix = F.width * F.height;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    (void)npix;
    // Slot q walks the image in 8x8 tiles so a wave64 traces one tile of
    // neighbouring camera rays; the path state stays indexed by pixel.
    const uint32_t tiles_x = (F.width + 7u) / 8u;
    const uint32_t tiles_y = (F.height + 7u) / 8u;
    const uint32_t nslots = tiles_x * tiles_y * 64u;
    for (uint32_t q0 = gid - threadIdx.x % 64u; q0 < nslots; q0 += stride) {
        const uint32_t qs = q0 + threadIdx.x % 64u;
        const uint32_t tile = qs >> 6, lane = qs & 63u;
        const uint32_t x = (tile % tiles_x) * 8u + (lane & 7u);
        const uint32_t y = (tile / tiles_x) * 8u + (lane >> 3);
        const bool valid = qs < nslots && x < F.width && y < F.height;
        const uint32_t slot = wave_append(&ctr[kCounterSlots - 1u], valid);
        if (!valid) continue;
        const uint32_t p = y * F.width + x;
        uint32_t seed = (s == 0u) ? pcg_hash(x * 1973u + y * 9277u + F.frame_index * 26699u) : ps.rng[p];
        float jx = randf(seed);
        float jy = randf(seed);
        jx = jx - 0.5f;
        jy = jy - 0.5f;
        float ux = ((((float)x + 0.5f) + jx) / (float)F.width) * 2.0f - 1.0f;
        float uy = ((((float)y + 0.5f) + jy) / (float)F.height) * 2.0f - 1.0f;
        uy = -uy;
        f3 rd = normalize((F.fwd + F.right * ((ux * F.aspect) * F.fovtan)) + F.up * (uy * F.fovtan));
        q.o[slot] = make_float4(F.pos.x, F.pos.y, F.pos.z, 0.0f);
        q.d[slot] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(p));
        ps.T[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        ps.L[p] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
        ps.rng[p] = seed;
    }
}

template <bool COUNT>
__global__ void __launch_bounds__(kBlock) k_extend(SceneView S, RayQueue q, const uint32_t* ctr, uint32_t b,
                                                   uint2* hits, uint32_t* spill, uint32_t spill_depth,
                                                   DeviceStats* st) {
    __shared__ uint32_t lds[kLdsDepth * kBlock];
    const uint32_t n = ctr[ctr_rays(b)];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    Stack stk = make_stack(lds, spill, spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    uint32_t rays = 0;
    for (uint32_t i = gid; i < n; i += stride) {
        const float4 o = q.o[i], d = q.d[i];
        HitRec h = trace_closest<COUNT>(S, xyz(o), xyz(d), stk, tc);
        hits[i] = make_uint2(__float_as_uint(h.t), h.code);
        rays++;
    }
    add_stat(b == 0u ? &st->primary : &st->extension, rays);
    if (COUNT) {
        add_stat(&st->ext_nodes, tc.nodes);
        add_stat(&st->ext_tris, tc.tris);
        add_stat(&st->ext_spheres, tc.spheres);
    }
}

__global__ void __launch_bounds__(kBlock) k_shade(SceneView S, FrameArgs F, uint32_t b, RayQueue qin,
                                                  RayQueue qout, uint32_t* ctr, const uint2* hits, PathState ps,
                                                  ShadowQueue sl, ShadowQueue se) {
    const uint32_t n = ctr[ctr_rays(b)];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool mis_on = F.mis_enabled != 0u;
    for (uint32_t i = gid; i < n; i += stride) {
        const float4 o4 = qin.o[i], d4 = qin.d[i];
        const uint32_t pix = __float_as_uint(d4.w);
        const f3 o = xyz(o4), d = xyz(d4);
        const uint2 hr = hits[i];
        const float4 Tv = ps.T[pix], Lv = ps.L[pix];
        uint32_t seed = ps.rng[pix];
        f3 T = xyz(Tv), L = xyz(Lv);
        float prev_pdf = Tv.w;
        uint32_t flags = __float_as_uint(Lv.w);
        bool skip_sky = (flags & 1u) != 0u, prev_lambert = (flags & 2u) != 0u;
        bool alive = false, emit_l = false, emit_e = false;
        f3 no{0.0f, 0.0f, 0.0f}, nd{0.0f, 1.0f, 0.0f};
        float4 lo{}, ld{}, la{}, lb{}, eo{}, ed{}, ea{}, eb{};
        float ew = 0.0f;
        if (hr.y == kMiss) {
            if (mis_on && prev_lambert) {
                // transmittance(ro, rd, 1e30) would re-trace this missed ray
                // and return exactly 1 (PathTrace.slang:6081-6084).
                float pe = pdf_env(S, d);
                float w = mis_balance(prev_pdf, pe);
                L = L + clamp_indirect(F, (T * sky_color(S, d)) * w);
            } else if (!skip_sky) {
                f3 c = T * sky_color(S, d);
                L = L + ((b == 0u) ? c : clamp_indirect(F, c));
            }
        } else {
            const float t = __uint_as_float(hr.x);
            const Surface sf = surface_at(S, o, d, HitRec{t, hr.y});
            const f3 hp = o + d * t;
            const float ndi = dot(sf.n, d);
            const f3 nf = (ndi < 0.0f) ? sf.n : -sf.n;
            if (any_pos(sf.emission)) {
                f3 c = T * sf.emission;
                L = L + ((b == 0u) ? c : clamp_indirect(F, c));
            }
            if (sf.mat == 0u) {
                if (S.light_count > 0u) {
                    LightSample ls = sample_light(S, F, hp, nf, seed);
                    float ndl = dot(nf, ls.dir);
                    if (ls.pdf > 0.0f && ndl > 0.0f) {
                        emit_l = true;
                        f3 so = hp + nf * 1e-3f, pre = T * ls.Le, brdf = sf.albedo / kPi;
                        lo = make_float4(so.x, so.y, so.z, fmaxf(ls.dist - 1e-3f, 0.0f));
                        ld = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, __uint_as_float(pix));
                        la = make_float4(pre.x, pre.y, pre.z, ndl);
                        lb = make_float4(brdf.x, brdf.y, brdf.z, ls.pdf);
                    }
                }
                if (S.env_present) {
                    float x0 = randf(seed);
                    float x1 = randf(seed);
                    EnvSample e = sample_env(S, x0, x1);
                    float ndl = dot(nf, e.dir);
                    if (e.pdf > 0.0f && ndl > 0.0f) {
                        emit_e = true;
                        f3 so = hp + nf * 1e-3f, pre = T * e.color, brdf = sf.albedo / kPi;
                        float pb = ndl * (1.0f / kPi);
                        bool fire = mis_on && (b + 1u) < F.max_bounces;
                        ew = fire ? mis_balance(e.pdf, pb) : 1.0f;
                        eo = make_float4(so.x, so.y, so.z, 1e30f);
                        ed = make_float4(e.dir.x, e.dir.y, e.dir.z, __uint_as_float(pix));
                        ea = make_float4(pre.x, pre.y, pre.z, ndl);
                        eb = make_float4(brdf.x, brdf.y, brdf.z, e.pdf);
                    }
                    skip_sky = !mis_on;
                } else {
                    // No env map: ambient albedo * skyColorBase(nf) (PathTrace.slang:6667),
                    // added after the light NEE term -> carried as a no-trace env record.
                    float tt = clampf(0.5f * (nf.y + 1.0f), 0.0f, 1.0f);
                    f3 amb = (T * sf.albedo) * lerp3(f3{0.5f, 0.55f, 0.62f}, f3{1.0f, 1.0f, 1.0f}, tt);
                    emit_e = true;
                    ed = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(pix));
                    ea = make_float4(amb.x, amb.y, amb.z, 0.0f);
                    eb = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                    skip_sky = true;
                }
                T = T * sf.albedo;
                no = hp + nf * 1e-3f;
                nd = cos_sample(nf, seed);
                if (mis_on && S.env_present) {
                    prev_pdf = fmaxf(0.0f, dot(nf, nd)) * (1.0f / kPi);
                    prev_lambert = true;
                } else {
                    prev_lambert = false;
                }
                alive = true;
            } else if (sf.mat == 1u) {
                f3 r = reflect(d, nf);
                f3 cs = cos_sample(nf, seed);
                r = normalize(r + (cs - nf) * sf.rough);
                T = T * sf.albedo;
                no = hp + nf * 1e-3f;
                nd = r;
                skip_sky = false;
                prev_lambert = false;
                alive = true;
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
                if (!take_refl) take_refl = randf(seed) < fr;
                if (take_refl) {
                    nd = refl;
                    no = hp + nf * 1e-3f;
                    skip_sky = false;
                    prev_lambert = false;
                } else {
                    nd = normalize(refr);
                    no = hp - nf * 1e-3f;
                }
                alive = true;
            }
            if (b >= 2u) {
                float p = clampf(maxc(T), 0.05f, 1.0f);
                if (randf(seed) > p) alive = false;
                else T = T / p;
            }
        }
        flags = (skip_sky ? 1u : 0u) | (prev_lambert ? 2u : 0u);
        ps.T[pix] = make_float4(T.x, T.y, T.z, prev_pdf);
        ps.L[pix] = make_float4(L.x, L.y, L.z, __uint_as_float(flags));
        ps.rng[pix] = seed;
        const bool cont = alive && (b + 1u) < F.max_bounces;
        uint32_t slot = wave_append(&ctr[ctr_rays(b + 1u)], cont);
        if (cont) {
            qout.o[slot] = make_float4(no.x, no.y, no.z, 0.0f);
            qout.d[slot] = make_float4(nd.x, nd.y, nd.z, __uint_as_float(pix));
        }
        slot = wave_append(&ctr[ctr_shadow(0, b)], emit_l);
        if (emit_l) { sl.o[slot] = lo; sl.d[slot] = ld; sl.a[slot] = la; sl.b[slot] = lb; }
        slot = wave_append(&ctr[ctr_shadow(1, b)], emit_e);
        if (emit_e) { se.o[slot] = eo; se.d[slot] = ed; se.a[slot] = ea; se.b[slot] = eb; se.w[slot] = ew; }
    }
}

// transmittance (PathTrace.slang:5089-5196): closest hits along the ray,
// refracting through dielectrics with a (1 - F) * albedo tint.
template <bool COUNT>
__device__ __forceinline__ f3 transmittance(const SceneView& S, const FrameArgs& F, f3 o, f3 d, float t_max,
                                            Stack& st, TraceCounters& tc, uint32_t& steps) {
    const int kmax = (int)F.refract_bounces;
    f3 t{1.0f, 1.0f, 1.0f};
    for (int i = 0; i <= kmax; ++i) {
        steps++;
        HitRec h = trace_closest<COUNT>(S, o, d, st, tc);
        if (h.code == kMiss || h.t >= t_max) return t;
        const Surface sf = surface_at(S, o, d, h);
        if (sf.mat != 2u) return f3{0.0f, 0.0f, 0.0f};
        if (i == kmax) return f3{0.0f, 0.0f, 0.0f};
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
        if (length(rf) < 1e-3f) return f3{0.0f, 0.0f, 0.0f};
        t = t * (sf.albedo * (1.0f - fr));
        float adv = h.t + 1e-3f;
        o = o + d * adv;
        d = normalize(rf);
        t_max = fmaxf(t_max - adv, 0.0f);
    }
    return f3{0.0f, 0.0f, 0.0f};
}

template <int KIND, bool COUNT>
__global__ void __launch_bounds__(kBlock) k_shadow(SceneView S, FrameArgs F, uint32_t b, ShadowQueue q,
                                                   const uint32_t* ctr, PathState ps, uint32_t* spill,
                                                   uint32_t spill_depth, DeviceStats* st) {
    __shared__ uint32_t lds[kLdsDepth * kBlock];
    const uint32_t n = ctr[ctr_shadow(KIND, b)];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    Stack stk = make_stack(lds, spill, spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    uint32_t steps = 0;
    for (uint32_t i = gid; i < n; i += stride) {
        const float4 o4 = q.o[i], d4 = q.d[i], a4 = q.a[i], b4 = q.b[i];
        const uint32_t pix = __float_as_uint(d4.w);
        const float4 Lv = ps.L[pix];
        f3 L = xyz(Lv);
        if (KIND == 1 && b4.w < 0.0f) {
            L = L + xyz(a4);   // ambient term of the no-env Lambert branch
        } else {
            f3 tr = transmittance<COUNT>(S, F, xyz(o4), xyz(d4), o4.w, stk, tc, steps);
            if (!any_pos(tr)) continue;
            f3 c;
            if (KIND == 0) {
                c = (((xyz(a4) * tr) * xyz(b4)) * a4.w) / b4.w;
                if (b > 0u) c = clamp_indirect(F, c);
            } else {
                c = ((((xyz(a4) * tr) * xyz(b4)) * a4.w) * q.w[i]) / b4.w;
                c = clamp_indirect(F, c);
            }
            L = L + c;
        }
        ps.L[pix] = make_float4(L.x, L.y, L.z, Lv.w);
    }
    add_stat(&st->shadow, steps);
    if (COUNT) {
        add_stat(&st->nodes, tc.nodes);
        add_stat(&st->tris, tc.tris);
        add_stat(&st->spheres, tc.spheres);
    }
}

__global__ void __launch_bounds__(kBlock) k_sample_end(FrameArgs F, PathState ps, float4* accum, uint32_t s) {
    const uint32_t npix = F.width * F.height;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = gid; p < npix; p += stride) {
        f3 fr = xyz(ps.L[p]);
        if (s > 0u) fr = xyz(ps.F[p]) + fr;
        if (s + 1u < F.spp) {
            ps.F[p] = make_float4(fr.x, fr.y, fr.z, 0.0f);
            continue;
        }
        fr = fr / (float)F.spp;
        float o[3] = {fr.x, fr.y, fr.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = (isnan(o[k]) || isinf(o[k])) ? 0.0f : fmaxf(o[k], 0.0f);
        float4 prev = F.reset_accum ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : accum[p];
        if (isnan(prev.x) || isinf(prev.x) || isnan(prev.y) || isinf(prev.y) || isnan(prev.z) || isinf(prev.z) ||
            isnan(prev.w) || isinf(prev.w))
            prev = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float ema = fminf(fmaxf(F.ema_alpha, 0.0f), 0.999f);
        float4 out;
        if (prev.w < 0.5f || ema <= 0.0f) {
            const float nn = prev.w + 1.0f;
            out = make_float4((prev.x * prev.w + o[0]) / nn, (prev.y * prev.w + o[1]) / nn,
                              (prev.z * prev.w + o[2]) / nn, nn);
        } else {
            out = make_float4(prev.x * ema + o[0] * (1.0f - ema), prev.y * ema + o[1] * (1.0f - ema),
                              prev.z * ema + o[2] * (1.0f - ema),
                              fminf(prev.w + 1.0f, 1.0f / fmaxf(1.0f - ema, 1e-3f)));
        }
        accum[p] = out;
    }
}

__global__ void __launch_bounds__(kBlock) k_trace_query(SceneView S, const float4* rays, uint32_t n, HitOut* out,
                                                        uint32_t* spill, uint32_t spill_depth) {
    __shared__ uint32_t lds[kLdsDepth * kBlock];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    Stack stk = make_stack(lds, spill, spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    for (uint32_t i = gid; i < n; i += stride) {
        const f3 o = xyz(rays[2 * i]), d = xyz(rays[2 * i + 1]);
        HitRec h = trace_closest<false>(S, o, d, stk, tc);
        HitOut r{};
        if (h.code != kMiss) {
            Surface sf = surface_at(S, o, d, h);
            r.t = h.t;
            r.n[0] = sf.n.x; r.n[1] = sf.n.y; r.n[2] = sf.n.z;
            r.hit = sf.kind;
            r.prim = sf.prim;
            r.mat = sf.mat;
            r.user_id = sf.user_id;
        } else {
            r.n[1] = 1.0f;
        }
        out[i] = r;
    }
}

// ---------------------------------------------------------------------------
// host launchers

FrameArgs make_frame_args(const float pf[4], const float fa[4], const float rt[4], const float up[4], uint32_t width,
                          uint32_t height, uint32_t frame_index, uint32_t spp, uint32_t max_bounces,
                          uint32_t reset_accum, uint32_t mis, uint32_t ltree, uint32_t refract, float clamp,
                          float ema) {
    FrameArgs F;
    F.pos = f3{pf[0], pf[1], pf[2]};
    F.fovtan = pf[3];
    F.fwd = f3{fa[0], fa[1], fa[2]};
    F.aspect = fa[3];
    F.right = f3{rt[0], rt[1], rt[2]};
    F.up = f3{up[0], up[1], up[2]};
    F.width = width;
    F.height = height;
    F.frame_index = frame_index;
    F.spp = spp;
    F.max_bounces = max_bounces;
    F.reset_accum = reset_accum;
    F.mis_enabled = mis;
    F.light_tree_enabled = ltree;
    F.refract_bounces = refract;
    F.firefly_clamp = clamp;
    F.ema_alpha = ema;
    return F;
}

uint32_t grid_for(uint32_t n) {
    uint32_t g = (n + kBlock - 1u) / kBlock;
    if (g < 1u) g = 1u;
    return g < kMaxGridBlocks ? g : kMaxGridBlocks;
}

template <bool COUNT>
static hipError_t frame_impl(const SceneView& S, const FrameArgs& F, Wavefront& W, float4* accum, hipStream_t stream,
                             hipEvent_t* ev, hipEvent_t accum_after) {
    const uint32_t npix = F.width * F.height;
    const uint32_t g = grid_for(npix);
    uint32_t k = 0;
    for (uint32_t s = 0; s < F.spp; ++s) {
        hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(kBlock), 0, stream, W.counters, npix);
        hipLaunchKernelGGL(k_raygen, dim3(g), dim3(kBlock), 0, stream, F, W.ps, W.rq[0], s, W.counters);
        for (uint32_t b = 0; b < F.max_bounces; ++b) {
            if (ev) (void)hipEventRecord(ev[2 * k], stream);
            hipLaunchKernelGGL(k_extend<COUNT>, dim3(g), dim3(kBlock), 0, stream, S, W.rq[b & 1u], W.counters, b,
                               W.hits, W.spill, W.spill_depth, W.stats);
            if (ev) (void)hipEventRecord(ev[2 * k + 1], stream);
            ++k;
            hipLaunchKernelGGL(k_shade, dim3(g), dim3(kBlock), 0, stream, S, F, b, W.rq[b & 1u],
                               W.rq[(b + 1u) & 1u], W.counters, W.hits, W.ps, W.sq[0], W.sq[1]);
            hipLaunchKernelGGL((k_shadow<0, COUNT>), dim3(g), dim3(kBlock), 0, stream, S, F, b, W.sq[0],
                               W.counters, W.ps, W.spill, W.spill_depth, W.stats);
            hipLaunchKernelGGL((k_shadow<1, COUNT>), dim3(g), dim3(kBlock), 0, stream, S, F, b, W.sq[1],
                               W.counters, W.ps, W.spill, W.spill_depth, W.stats);
        }
        if (s + 1u == F.spp && accum_after) (void)hipStreamWaitEvent(stream, accum_after, 0);
        hipLaunchKernelGGL(k_sample_end, dim3(g), dim3(kBlock), 0, stream, F, W.ps, accum, s);
    }
    return hipGetLastError();
}

hipError_t launch_frame(const SceneView& S, const FrameArgs& F, Wavefront& W, float4* accum, hipStream_t stream,
                        hipEvent_t* ev, hipEvent_t accum_after) {
    return frame_impl<false>(S, F, W, accum, stream, ev, accum_after);
}

hipError_t launch_frame_counting(const SceneView& S, const FrameArgs& F, Wavefront& W, float4* accum,
                                 hipStream_t stream) {
    return frame_impl<true>(S, F, W, accum, stream, nullptr, nullptr);
}

hipError_t launch_trace_query(const SceneView& S, const float4* rays, uint32_t n, HitOut* out, Wavefront& W,
                              hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_query, dim3(grid_for(n)), dim3(kBlock), 0, stream, S, rays, n, out, W.spill,
                       W.spill_depth);
    return hipGetLastError();
}

}  // namespace dmt
