// Adjoint (path replay) kernel: see adjoint.h.
#include "adjoint.h"
#include "dual.h"

namespace dmt {

namespace {

__device__ __forceinline__ uint32_t argmax3(f3 v) {
    const float m = maxc(v);
    return (v.x == m) ? 0u : ((v.y == m) ? 1u : 2u);
}

__device__ __forceinline__ float comp(f3 v, uint32_t k) { return k == 0u ? v.x : (k == 1u ? v.y : v.z); }

__device__ __forceinline__ bool finite3(f3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

__device__ __forceinline__ void add_grad(float* g, uint32_t i, float v) {
    if (v != 0.0f && isfinite(v)) atomicAdd(g + i, v);
}

// Path-replay sink: forms the radiance terms exactly as the renderer does
// (shadow queries traced at once, in emission order) and scatters each
// term's derivative.  REC: records the sample's metal vertices on analytic
// primitives (at most kAdjMaxVerts) for the roughness pass.  SDF: the
// queries see the scene's SDF clusters.
template <bool FEAT, bool REC = false, bool SDF = false>
struct AdjointSink : NoGradientHooks {
    const SceneView& S;
    const FrameArgs& F;
    Stack& st;
    TraceCounters& tc;
    float* G;                          // gradient accumulator (LDS or global)
    f3 delta;                          // dLoss/dI of the pixel / spp
    uint32_t n;                        // albedo vertices so far
    uint32_t prim[kAdjMaxVerts];
    f3 acc[kAdjMaxVerts];
    float4* rec;                       // REC: the record list and its length
    uint32_t* rec_count;
    uint32_t metal;                    // REC: metal vertices recorded for this sample

    // A metal vertex of an analytic primitive: its record, appended with one
    // counter atomic per wave (the lanes at such a vertex take consecutive
    // records).  Samples whose dLoss/dI is zero have no roughness gradient.
    __device__ __forceinline__ void on_vertex(uint32_t b, const Surface& sf, f3 o, f3 d, HitRec h,
                                              const PathVertexState& P) {
        if (!REC || sf.mat != 1u || sf.kind != 1u || metal >= kAdjMaxVerts) return;
        if (delta.x == 0.0f && delta.y == 0.0f && delta.z == 0.0f) return;
        metal++;
        const unsigned long long m = __ballot(1);
        const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
        uint32_t base = 0u;
        if (__lane_id() == leader) base = atomicAdd(rec_count, (uint32_t)__popcll(m));
        base = __shfl(base, (int)leader);
        const uint32_t i = base + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
        float4* r = rec + (size_t)i * kAdjRecFloat4;
        const uint32_t bits = b | (P.skip_sky ? 256u : 0u) | (P.prev_lambert ? 512u : 0u);
        r[0] = make_float4(o.x, o.y, o.z, h.t);
        r[1] = make_float4(d.x, d.y, d.z, __uint_as_float(h.code));
        r[2] = make_float4(P.T_.x, P.T_.y, P.T_.z, P.prev_pdf);
        r[3] = make_float4(__uint_as_float(P.seed), __uint_as_float(sf.prim), __uint_as_float(bits), P.lambda);
        r[4] = make_float4(delta.x, delta.y, delta.z, P.time);
    }

    // delta . dc / dlog(albedo_ch), per channel ch.
    __device__ __forceinline__ f3 log_derivative(f3 v, f3 c, bool may_clamp) const {
        f3 q = delta * c;
        const bool clamped = may_clamp && F.firefly_clamp > 0.0f && maxc(v) > F.firefly_clamp;
        if (clamped) {
            const float dc = dot(delta, c);
            const uint32_t k = argmax3(v);
            if (k == 0u) q.x = q.x - dc;
            else if (k == 1u) q.y = q.y - dc;
            else q.z = q.z - dc;
        }
        return q;
    }

    __device__ __forceinline__ void add_prefix(f3 q) {
#pragma unroll
        for (uint32_t i = 0; i < kAdjMaxVerts; ++i)
            if (i < n) acc[i] = acc[i] + q;
    }

    __device__ __forceinline__ void on_radiance(f3 v, f3 c, bool may_clamp) {
        if (n > 0u) add_prefix(log_derivative(v, c, may_clamp));
    }

    __device__ __forceinline__ void on_emission(const Surface& sf, f3 T, f3 v, f3 c, bool may_clamp) {
        if (sf.kind != 1u) return;
        const float M = F.firefly_clamp;
        const float m = maxc(v);
        const bool clamped = may_clamp && M > 0.0f && m > M;
        const float s = clamped ? (M / m) : 1.0f;
        f3 g = (delta * T) * s;
        if (clamped) {
            const uint32_t k = argmax3(v);
            const float corr = (dot(delta, c) * comp(T, k)) / m;
            if (k == 0u) g.x = g.x - corr;
            else if (k == 1u) g.y = g.y - corr;
            else g.z = g.z - corr;
        }
        const uint32_t base = sf.prim * kMaterialFloats + 4u;
        add_grad(G, base + 0u, g.x);
        add_grad(G, base + 1u, g.y);
        add_grad(G, base + 2u, g.z);
    }

    __device__ __forceinline__ void on_albedo_vertex(const Surface& sf) {
        if (sf.kind != 1u || n >= kAdjMaxVerts) return;
#pragma unroll
        for (uint32_t i = 0; i < kAdjMaxVerts; ++i) {
            if (i == n) {
                prim[i] = sf.prim;
                acc[i] = f3{0.0f, 0.0f, 0.0f};
            }
        }
        n++;
    }

    __device__ __forceinline__ void emit(int kind, const ShadowRec& r, PathVertexState& P) {
        if (kind == 1 && r.pdf < 0.0f) {   // ambient term: no query, never clamped
            P.L = P.L + r.pre;
            add_prefix(delta * r.pre);
            return;
        }
        // transmittance, recording the dielectrics the query passes through
        uint32_t diel[kAdjMaxDielectric];
        uint32_t nd = 0;
        f3 o = r.o, d = r.d, tr{1.0f, 1.0f, 1.0f};
        float t_max = r.t_max;
        bool done = false;
        for (uint32_t i = 0; i <= F.refract_bounces && !done; ++i) {
            const HitRec h = trace_closest<false, FEAT, SDF>(S, o, d, st, tc, P.time);
            done = transmit_step<SDF>(S, F, h, i, o, d, t_max, tr, P.time);
#pragma unroll
            for (uint32_t j = 0; j < kAdjMaxDielectric; ++j)
                if (!done && j == nd) diel[j] = h.code;
            if (!done) nd++;
        }
        if (!done) tr = f3{0.0f, 0.0f, 0.0f};
        if (!any_pos(tr)) return;
        f3 v;
        if (kind == 0) v = (((r.pre * tr) * r.brdf) * r.ndl) / r.pdf;
        else v = ((((r.pre * tr) * r.brdf) * r.ndl) * r.w) / r.pdf;
        const f3 c = r.clamp ? clamp_indirect(F, v) : v;
        P.L = P.L + c;
        const f3 q = log_derivative(v, c, r.clamp);
        add_prefix(q);
#pragma unroll
        for (uint32_t j = 0; j < kAdjMaxDielectric; ++j) {
            if (j < nd && !(diel[j] & (kMeshBit | kSdfBit))) {   // dielectric SDF clusters have no parameters
                const uint32_t pid = diel[j];
                const float4 a = S.prims[(size_t)pid * 7u + 1u];
                const uint32_t base = pid * kMaterialFloats;
                if (a.x != 0.0f) add_grad(G, base + 0u, q.x / a.x);
                if (a.y != 0.0f) add_grad(G, base + 1u, q.y / a.y);
                if (a.z != 0.0f) add_grad(G, base + 2u, q.z / a.z);
            }
        }
    }

    __device__ __forceinline__ void flush() {
#pragma unroll
        for (uint32_t i = 0; i < kAdjMaxVerts; ++i) {
            if (i < n && finite3(acc[i])) {
                const float4 a = S.prims[(size_t)prim[i] * 7u + 1u];
                const uint32_t base = prim[i] * kMaterialFloats;
                if (a.x != 0.0f) add_grad(G, base + 0u, acc[i].x / a.x);
                if (a.y != 0.0f) add_grad(G, base + 1u, acc[i].y / a.y);
                if (a.z != 0.0f) add_grad(G, base + 2u, acc[i].z / a.z);
            }
        }
    }
};

// SDF (implies FEAT): the scene has SDF clusters.  Their surfaces carry no
// parameters; the replay traces and shades them as the renderer does.
template <bool LDS, bool FEAT, bool REC, bool SDF = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kAdjWavesPerSimd, kAdjWavesPerSimd)))
k_adjoint(AdjointArgs A, uint32_t p0, uint32_t p1) {
    __shared__ uint32_t lds_stack[kLdsDepth * kBlock];
    __shared__ float lds_grad[LDS ? kAdjLdsPrims * kMaterialFloats : 1u];
    const uint32_t nparam = A.prim_count * kMaterialFloats;
    if (LDS) {
        for (uint32_t i = threadIdx.x; i < nparam; i += blockDim.x) lds_grad[i] = 0.0f;
        __syncthreads();
    }
    float* G = LDS ? lds_grad : A.grads;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    Stack stk = stack_at(lds_stack + threadIdx.x, kBlock, kLdsDepth, A.spill + (size_t)gid * A.spill_depth, A.spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    const FrameArgs& F = A.F;
    for (uint32_t p = p0 + gid; p < p1; p += stride) {
        const uint32_t x = p % F.width, y = p / F.width;
        const f3 delta = xyz(A.dimage[p]) / (float)F.spp;
        uint32_t seed = pixel_seed(F, x, y, F.frame_index);
        for (uint32_t s = 0; s < F.spp; ++s) {
            f3 o, d;
            float time, lambda;
            camera_sample(F, A.X, x, y, s, seed, o, d, time, lambda);
            PathVertexState P{f3{1.0f, 1.0f, 1.0f}, f3{0.0f, 0.0f, 0.0f}, 0.0f, false, false, seed, time, lambda};
            // a spectral sample enters the frame value weighted by its wavelength's response
            const f3 ds = (lambda >= 0.0f) ? delta * spectral_weight(lambda) : delta;
            AdjointSink<FEAT, REC, SDF> sink{{}, A.S, F, stk, tc, G, ds, 0u, {}, {}, A.rec, A.rec_count, 0u};
            for (uint32_t b = 0; b < F.max_bounces; ++b) {
                const HitRec h = trace_closest<false, FEAT, SDF>(A.S, o, d, stk, tc, time);
                VertexResult R;
                shade_vertex<false, SDF>(A.S, F, b, o, d, h, P, R, sink);
                if (!(R.alive && (b + 1u) < F.max_bounces)) break;
                o = R.no;
                d = R.nd;
            }
            seed = P.seed;
            if (finite3(P.L)) sink.flush();
        }
    }
    if (LDS) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nparam; i += blockDim.x) {
            const float v = lds_grad[i];
            if (v != 0.0f) atomicAdd(A.grads + i, v);
        }
    }
}

// ---------------------------------------------------------------------------
// Roughness: forward-mode tangents.  A metal vertex's roughness bends its
// reflected ray, which moves every later hit, normal, light sample and shadow
// query; terms before the vertex do not depend on it.  k_adjoint recorded
// each metal vertex with the ray that reached it, its hit and the path state
// it arrived with (the state the renderer has there: the sink does not touch
// the random stream).  Each record is continued here in dual numbers with
// d(roughness) = 1 at that vertex, and delta . dL/d(roughness) goes to the
// vertex's primitive.  The hit set, the material of every hit and every
// sampling decision are held fixed, so this is the derivative away from
// visibility boundaries.  The records are a flat list, so the grid's lanes
// take one vertex each whatever pixel or sample it belongs to.

// Dual vertices: roughness derivative seeded at bounce seed_b; shadow queries
// traced at once along the dual ray.
template <bool FEAT, bool SDF = false>
struct TangentSink : NoGradientHooks {
    const SceneView& S;
    const FrameArgs& F;
    Stack& st;
    TraceCounters& tc;
    uint32_t seed_b;
    __device__ __forceinline__ void on_surface(uint32_t b, SurfaceT<df>& sf) {
        if (b == seed_b) sf.rough.d = 1.0f;
    }
    __device__ __forceinline__ void emit(int kind, const ShadowRecT<df>& r, PathVertexStateT<df>& P) {
        if (kind == 1 && r.pdf < 0.0f) {
            P.L = P.L + r.pre;
            return;
        }
        d3 o = r.o, d = r.d;
        d3 tr = to<df>(f3{1.0f, 1.0f, 1.0f});
        df t_max = r.t_max;
        bool done = false;
        for (uint32_t i = 0; i <= F.refract_bounces && !done; ++i) {
            const HitRec h = trace_closest<false, FEAT, SDF>(S, value3(o), value3(d), st, tc, P.time);
            done = transmit_step<SDF>(S, F, h, i, o, d, t_max, tr, P.time);
        }
        if (!done || !any_pos(tr)) return;
        P.L = P.L + shadow_value(F, kind, r.pre, r.ndl, r.brdf, r.pdf, r.w, r.clamp, tr);
    }
};

template <bool LDS, bool FEAT, bool SDF = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kAdjWavesPerSimd, kAdjWavesPerSimd)))
k_adjoint_dual(AdjointArgs A) {
    __shared__ uint32_t lds_stack[kLdsDepth * kBlock];
    __shared__ float lds_grad[LDS ? kAdjLdsPrims : 1u];
    if (LDS) {
        for (uint32_t i = threadIdx.x; i < A.prim_count; i += blockDim.x) lds_grad[i] = 0.0f;
        __syncthreads();
    }
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    Stack stk = stack_at(lds_stack + threadIdx.x, kBlock, kLdsDepth, A.spill + (size_t)gid * A.spill_depth, A.spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    const FrameArgs& F = A.F;
    const uint32_t n = min(*A.rec_count, A.rec_cap);
    for (uint32_t e = gid; e < n; e += stride) {
        const float4* r = A.rec + (size_t)e * kAdjRecFloat4;
        const float4 e0 = r[0], e1 = r[1], e2 = r[2], e3 = r[3], e4 = r[4];
        const uint32_t bits = __float_as_uint(e3.z);
        const uint32_t bj = bits & 0xFFu;
        const uint32_t pj = __float_as_uint(e3.y);
        d3 od = to<df>(xyz(e0)), dd = to<df>(xyz(e1));
        HitRec h{e0.w, __float_as_uint(e1.w)};
        PathVertexStateT<df> PD{to<df>(xyz(e2)), to<df>(f3{0.0f, 0.0f, 0.0f}), df(e2.w), (bits & 256u) != 0u,
                                (bits & 512u) != 0u, __float_as_uint(e3.x), e4.w, e3.w};
        TangentSink<FEAT, SDF> ts{{}, A.S, F, stk, tc, bj};
        for (uint32_t b = bj; b < F.max_bounces; ++b) {
            if (b != bj) h = trace_closest<false, FEAT, SDF>(A.S, value3(od), value3(dd), stk, tc, PD.time);
            VertexResultT<df> R;
            shade_vertex<false, SDF>(A.S, F, b, od, dd, h, PD, R, ts);
            if (!(R.alive && (b + 1u) < F.max_bounces)) break;
            od = R.no;
            dd = R.nd;
        }
        const float g = dot(xyz(e4), deriv3(PD.L));
        if (LDS) add_grad(lds_grad, pj, g);
        else add_grad(A.grads, pj * kMaterialFloats + 3u, g);
    }
    if (LDS) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < A.prim_count; i += blockDim.x) {
            const float v = lds_grad[i];
            if (v != 0.0f) atomicAdd(A.grads + i * kMaterialFloats + 3u, v);
        }
    }
}

__global__ void k_fill_zero(float* p, uint32_t n) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0f;
}

// One chunk of pixels [p0, p1): the replay (recording metal vertices when
// rough) and the roughness pass over the records, in variant <FT, SDF>.
template <bool FT, bool SDF>
void adjoint_pass(const AdjointArgs& A, bool lds, bool rough, uint32_t blocks, uint32_t dual_blocks, uint32_t p0,
                  uint32_t p1, hipStream_t stream) {
    if (rough) {
        if (lds) hipLaunchKernelGGL((k_adjoint<true, FT, true, SDF>), dim3(blocks), dim3(kBlock), 0, stream, A, p0, p1);
        else hipLaunchKernelGGL((k_adjoint<false, FT, true, SDF>), dim3(blocks), dim3(kBlock), 0, stream, A, p0, p1);
        if (lds) hipLaunchKernelGGL((k_adjoint_dual<true, FT, SDF>), dim3(dual_blocks), dim3(kBlock), 0, stream, A);
        else hipLaunchKernelGGL((k_adjoint_dual<false, FT, SDF>), dim3(dual_blocks), dim3(kBlock), 0, stream, A);
    } else {
        if (lds) hipLaunchKernelGGL((k_adjoint<true, FT, false, SDF>), dim3(blocks), dim3(kBlock), 0, stream, A, p0, p1);
        else hipLaunchKernelGGL((k_adjoint<false, FT, false, SDF>), dim3(blocks), dim3(kBlock), 0, stream, A, p0, p1);
    }
}

}  // namespace

hipError_t launch_adjoint(const AdjointArgs& A, hipStream_t stream) {
    const uint32_t nparam = A.prim_count * kMaterialFloats;
    if (nparam) hipLaunchKernelGGL(k_fill_zero, dim3(min(1024u, (nparam + 255u) / 256u)), dim3(256), 0, stream, A.grads, nparam);
    const uint32_t npix = A.F.width * A.F.height;
    if (npix == 0 || A.F.max_bounces == 0 || nparam == 0) return hipGetLastError();
    const bool lds = A.prim_count <= kAdjLdsPrims;
    const bool rough = A.metal_prims != 0u;
    const uint32_t chunk = rough ? A.chunk_px : npix;
    if (chunk == 0u) return hipErrorInvalidValue;
    // the records' pass: a resident grid (one workgroup per SIMD and wave
    // slot), every lane on the next record until the list is done
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const uint32_t dual_blocks = min(kMaxGridBlocks, (uint32_t)(cus > 0 ? cus : 1) * kAdjWavesPerSimd);
    for (uint32_t p0 = 0; p0 < npix; p0 += chunk) {
        const uint32_t p1 = npix - p0 > chunk ? p0 + chunk : npix;
        const uint32_t blocks = min(kMaxGridBlocks, (p1 - p0 + kBlock - 1u) / kBlock);
        if (rough) hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(64), 0, stream, (float*)A.rec_count, 1u);
        if (!A.features) adjoint_pass<false, false>(A, lds, rough, blocks, dual_blocks, p0, p1, stream);
        else if (A.X.ext & kExtSdf) adjoint_pass<true, true>(A, lds, rough, blocks, dual_blocks, p0, p1, stream);
        else adjoint_pass<true, false>(A, lds, rough, blocks, dual_blocks, p0, p1, stream);
    }
    return hipGetLastError();
}

}  // namespace dmt
