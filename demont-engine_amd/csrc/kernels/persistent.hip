// Persistent batch kernel (see persistent.h).
#include "persistent.h"

namespace dmt {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kEnded = 256u;         // pending bit: the path ends after this vertex

// lane states
constexpr uint32_t kLaneIdle = 0u;    // no path
constexpr uint32_t kLaneWalk = 1u;    // a closest-hit query in progress
constexpr uint32_t kLaneReady = 2u;   // the query is done; its hit waits for the wave's shading step
#ifndef DMT_PERSIST_SHADE_LANES
#define DMT_PERSIST_SHADE_LANES 40u   // as k_tail's (the 64-scene batch: 0.4 % faster than 32)
#endif
constexpr uint32_t kPersistShadeLanes = DMT_PERSIST_SHADE_LANES;    // shade once this many lanes are ready
constexpr uint32_t kPersistRefillLanes = 16u;   // take pixels once this many lanes are idle

// Shadow records of one vertex go to this thread's scratch rows; bit k of
// `pending` marks record k, bit k + 4 its clampIndirect flag.
struct ScratchSink : NoGradientHooks {
    float4* rec;
    uint32_t threads, tid;
    uint32_t pending;
    __device__ __forceinline__ void emit(int kind, const ShadowRec& r, PathVertexState&) {
        float4* row = rec + ((size_t)kind * kRecRows) * threads + tid;
        row[0] = make_float4(r.o.x, r.o.y, r.o.z, r.t_max);
        row[(size_t)threads] = make_float4(r.d.x, r.d.y, r.d.z, r.w);
        row[2 * (size_t)threads] = make_float4(r.pre.x, r.pre.y, r.pre.z, r.ndl);
        row[3 * (size_t)threads] = make_float4(r.brdf.x, r.brdf.y, r.brdf.z, r.pdf);
        pending |= (1u << kind) | (r.clamp ? (16u << kind) : 0u);
    }
};

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void add_stat(unsigned long long* dst, uint32_t v) {
    uint32_t s = wave_sum(v);
    if (__lane_id() == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// The scene table is constant while the kernel runs.  Read through the
// constant address space, a wave-uniform scene's fields are scalar loads
// that the compiler may issue again where they are used (like kernel
// arguments) instead of holding them in registers across the loop.
template <class T> using ConstPtr = __attribute__((address_space(4))) const T*;

template <class T> __device__ __forceinline__ T const_copy(ConstPtr<T> src) {
    static_assert(sizeof(T) % 4u == 0u, "dword-sized struct");
    T out;
    const ConstPtr<uint32_t> s = (ConstPtr<uint32_t>)src;
    uint32_t* d = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4u; ++i) d[i] = s[i];
    return out;
}

// State of one lane's path.  The fields used at every query stay in
// registers; the rest sits in an LDS column of the thread (`Cold`), so the
// registers of the traversal and of the vertex shading are not held by it.
struct Lane {
    bool active;
    uint32_t mode;        // 0 extension ray, 1 light-NEE shadow ray, 2 env-NEE shadow ray
    f3 o, d;              // current ray
    PathVertexState P;
    uint32_t b;           // bounce
    uint32_t pending;     // shadow records of the vertex, their clamp flags, kEnded
};

struct Cold {
    float* base;          // this thread's column: field f at base[f * kBlock]
    // kXY: x | y << 16 (frames are at most 65535 pixels wide and high)
    enum : uint32_t { kSum = 0, kXY = 3, kS = 4, kTr = 5, kTmax = 8, kStep = 9, kNo = 10, kNd = 13, kRows = 16 };
    __device__ __forceinline__ float get(uint32_t f) const { return base[f * kBlock]; }
    __device__ __forceinline__ void set(uint32_t f, float v) { base[f * kBlock] = v; }
    __device__ __forceinline__ uint32_t getu(uint32_t f) const { return __float_as_uint(get(f)); }
    __device__ __forceinline__ void setu(uint32_t f, uint32_t v) { set(f, __uint_as_float(v)); }
    __device__ __forceinline__ f3 get3(uint32_t f) const { return f3{get(f), get(f + 1), get(f + 2)}; }
    __device__ __forceinline__ void set3(uint32_t f, f3 v) { set(f, v.x); set(f + 1, v.y); set(f + 2, v.z); }
};

// Sample s of pixel (x, y); FEAT: with the frame's camera-sample features.
// A sample starts with a zero ray cone and, in WATER frames, in the camera's
// medium.
template <bool FEAT, bool WATER = false>
__device__ __forceinline__ void start_sample(const FrameArgs& F, const FeatureArgs& X, Lane& l, uint32_t x,
                                             uint32_t y, uint32_t s) {
    if (FEAT) {
        camera_sample(F, X, x, y, s, l.P.seed, l.o, l.d, l.P.time, l.P.lambda);
    } else {
        l.d = camera_ray(F, x, y, l.P.seed);
        l.o = F.pos;
    }
    l.P.T_ = f3{1.0f, 1.0f, 1.0f};
    l.P.L = f3{0.0f, 0.0f, 0.0f};
    l.P.prev_pdf = 0.0f;
    l.P.skip_sky = false;
    l.P.prev_lambert = false;
    l.P.cone = 0.0f;
    if (WATER) l.P.in_water = X.cam_in_water != 0u;
    l.b = 0u;
    l.mode = 0u;
}

// A shadow record's query, in the path's medium (bit 31 of the step field):
// a Lambert vertex, whose NEE records these are, does not change it.
template <bool WATER = false>
__device__ __forceinline__ void start_shadow(const float4* rec, uint32_t threads, uint32_t tid, uint32_t k, Lane& l,
                                             Cold& c) {
    const float4* row = rec + ((size_t)k * kRecRows) * threads + tid;
    const float4 o4 = row[0], d4 = row[(size_t)threads];
    l.o = xyz(o4);
    l.d = xyz(d4);
    c.set(Cold::kTmax, o4.w);
    c.set3(Cold::kTr, f3{1.0f, 1.0f, 1.0f});
    c.setu(Cold::kStep, (WATER && l.P.in_water) ? 0x80000000u : 0u);
    l.mode = k + 1u;
}

// After the vertex's shading or one of its shadow records: start the next
// record in the order the vertex's terms are summed -- light, then env NEE
// and the HDRI cluster NEE with an env map; light, sun, moon, then the
// ambient term without one -- or close the vertex.  Records 2 and 3 (sky
// lights) exist only in SKY frames.
template <bool FEAT, bool SKY = false, bool WATER = false>
__device__ __forceinline__ void advance(const FrameArgs& F, const FeatureArgs& X, const float4* rec, uint32_t threads,
                                        uint32_t tid, float4* accum, bool reset, Lane& l, Cold& c) {
    if (l.pending & 1u) {
        start_shadow<WATER>(rec, threads, tid, 0u, l, c);
        return;
    }
    if (!SKY) {
        if (l.pending & 2u) {
            const float4* row = rec + ((size_t)1u * kRecRows) * threads + tid;
            const float4 b4 = row[3 * (size_t)threads];
            if (b4.w < 0.0f) {
                // ambient term of the no-env Lambert branch: added without a query
                l.P.L = l.P.L + xyz(row[2 * (size_t)threads]);
                l.pending &= ~2u;
            } else {
                start_shadow<WATER>(rec, threads, tid, 1u, l, c);
                return;
            }
        }
    } else {
        bool ambient = false;
        if (l.pending & 2u) {
            const float4* row = rec + ((size_t)1u * kRecRows) * threads + tid;
            ambient = row[3 * (size_t)threads].w < 0.0f;   // the no-env ambient term: after the sun and moon
            if (!ambient) {
                start_shadow<WATER>(rec, threads, tid, 1u, l, c);
                return;
            }
        }
        if (l.pending & 4u) {
            start_shadow<WATER>(rec, threads, tid, 2u, l, c);
            return;
        }
        if (l.pending & 8u) {
            start_shadow<WATER>(rec, threads, tid, 3u, l, c);
            return;
        }
        if (ambient) {
            const float4* row = rec + ((size_t)1u * kRecRows) * threads + tid;
            l.P.L = l.P.L + xyz(row[2 * (size_t)threads]);
            l.pending &= ~2u;
        }
    }
    if (!(l.pending & kEnded)) {
        l.o = c.get3(Cold::kNo);
        l.d = c.get3(Cold::kNd);
        l.b++;
        l.mode = 0u;
        return;
    }
    // the sample is done
    const uint32_t s = c.getu(Cold::kS);
    f3 sum;
    if (FEAT) {   // the hero wavelength's response (PathTrace.slang:7693-7700)
        const f3 v = (l.P.lambda >= 0.0f) ? l.P.L * spectral_weight(l.P.lambda) : l.P.L;
        sum = ((s == 0u) ? f3{0.0f, 0.0f, 0.0f} : c.get3(Cold::kSum)) + v;
    } else {
        sum = (s == 0u) ? l.P.L : c.get3(Cold::kSum) + l.P.L;
    }
    const uint32_t xy = c.getu(Cold::kXY), x = xy & 0xFFFFu, y = xy >> 16;
    if (s + 1u < F.spp) {
        c.set3(Cold::kSum, sum);
        c.setu(Cold::kS, s + 1u);
        start_sample<FEAT, WATER>(F, X, l, x, y, s + 1u);   // the next sample continues the pixel's stream
        return;
    }
    const uint32_t p = y * F.width + x;
    const f3 v = frame_value(F, sum, FEAT && X.spectral != 0u);
    accum[p] = accumulate(F, reset ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : accum[p], reset, v);
    l.active = false;
}

// Workgroups per CU that the LDS of one (stack + cold column) allows, and
// the occupancy the register budget is set for.
template <uint32_t DEPTH> constexpr uint32_t persistent_waves() { return (160u * 1024u) / ((DEPTH + Cold::kRows) * kBlock * 4u); }

// FEAT: some scene's frame uses camera-sample features (thin lens, shutter
// time, hero wavelength, a translated mesh); each scene's FeatureArgs say
// which, and the lane carries the sample's shutter time and wavelength.
// EXT (implies FEAT): kExtSky, some scene's frame has sky lights (four record
// kinds); kExtSdf, some scene has SDF clusters; kExtPbr, some scene's frame
// samples the PBR atlas; kExtWater, some scene has water or a camera in it.
template <uint32_t DEPTH, bool FEAT, uint32_t EXT = 0u>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(persistent_waves<DEPTH>(), persistent_waves<DEPTH>()))) k_persistent(BatchArgs A) {
    __shared__ uint32_t lds[DEPTH * kBlock];
    __shared__ float cold[Cold::kRows * kBlock];
    Cold c{cold + threadIdx.x};
    const uint32_t tid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    Stack stk = stack_at(lds + threadIdx.x, kBlock, DEPTH, A.spill + (size_t)tid * A.spill_depth, A.spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    uint32_t n_primary = 0, n_ext = 0, n_shadow = 0;
    const uint32_t rel = *A.frame;

    // wave-uniform work position
    const uint32_t xcd = blockIdx.x % kXcds;
    uint32_t qoff = 0u, qpos = A.queue_begin[xcd];
    bool open = false;                 // the current scene may have tiles left
    uint32_t sc = kNone, pool_lo = 0u, pool_hi = 0u;
    const ConstPtr<BatchScene> table = (ConstPtr<BatchScene>)A.scenes;
    ConstPtr<BatchScene> bs = table;
    uint32_t frame_index = 0u;
    float4* accum = nullptr;
    uint32_t tiles_x = 1u, tiles = 0u;
    bool reset = false;

    Lane l{};
    l.active = false;
    const uint32_t tw = A.tile_w, tsh = A.tile_shift;
    // Each lane advances its query one node or leaf per loop step (Walk); a
    // lane whose query is done waits (kLaneReady) until enough of the wave is
    // done too, and the wave then shades those lanes together.
    Walk w{};
    HitRec h{0.0f, kMiss};
    uint32_t state = kLaneIdle;
    for (;;) {
        // ---- hand pixels to idle lanes (wave-uniform control flow) ----
        // The wave may switch scene (bs) in this loop, so the scene's state
        // is read through bs where it is used, never held across a switch.
        // It takes pixels once enough lanes are idle, or none is walking.
        const bool take = (uint32_t)__popcll(__ballot(!l.active)) >= kPersistRefillLanes ||
                          __ballot(state == kLaneWalk) == 0ull;
        for (; take;) {
            const unsigned long long idle = __ballot(!l.active);
            if (idle == 0ull) break;
            if (pool_lo < pool_hi) {
                const FrameArgs F = const_copy(&bs->F);
                const SceneView S = const_copy(&bs->S);
                const FeatureArgs X = FEAT ? const_copy(&bs->X) : FeatureArgs{};
                const uint32_t ord = pool_lo + (uint32_t)__popcll(idle & below);
                if (!l.active && ord < pool_hi) {
                    const uint32_t tile = ord >> 6, k = ord & 63u;
                    const uint32_t x = (tile % tiles_x) * tw + (k & (tw - 1u));
                    const uint32_t y = (tile / tiles_x) * (64u >> tsh) + (k >> tsh);
                    if (x < F.width && y < F.height) {
                        l.active = true;
                        c.setu(Cold::kXY, x | (y << 16));
                        c.setu(Cold::kS, 0u);
                        l.P.seed = pixel_seed(F, x, y, frame_index);
                        start_sample<FEAT, (EXT & kExtWater) != 0u>(F, X, l, x, y, 0u);
                        state = kLaneWalk;
                        if (walk_begin<(EXT & kExtSdf) != 0u>(S, l.o, l.d, w)) {
                            h = final_hit(w.ab, w.tb);
                            state = kLaneReady;
                        }
                    }
                }
                pool_lo = min(pool_lo + (uint32_t)__popcll(idle), pool_hi);
                continue;
            }
            if (open) {
                uint32_t start = 0u;
                if (lane == 0u) start = atomicAdd(&A.work[sc], A.fetch);
                start = uniform(__shfl(start, 0));
                if (start < tiles) {
                    pool_lo = start * 64u;
                    pool_hi = min(start + A.fetch, tiles) * 64u;
                    continue;
                }
                open = false;
            }
            if (__ballot(l.active) != 0ull) break;   // finish this scene's paths first
            // next scene with tiles left: this XCD's queue, then the others
            sc = kNone;
            while (qoff < kXcds) {
                const uint32_t q = (xcd + qoff) % kXcds;
                if (qpos < A.queue_begin[q + 1u]) {
                    const uint32_t cand = uniform(A.order[qpos++]);
                    const uint32_t handed = __hip_atomic_load(&A.work[cand], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (uniform(handed) < table[cand].tiles) {
                        sc = cand;
                        break;
                    }
                } else {
                    ++qoff;
                    qpos = A.queue_begin[(xcd + qoff) % kXcds];
                }
            }
            if (sc == kNone) break;   // no work left anywhere
            bs = table + sc;
            frame_index = bs->F.frame_index + rel;
            reset = rel == 0u && bs->F.reset_accum != 0u;
            accum = bs->accum;
            tiles_x = bs->tiles_x;
            tiles = bs->tiles;
            open = true;
            pool_lo = pool_hi = 0u;
        }
        const unsigned long long ready = __ballot(state == kLaneReady);
        const unsigned long long walk = __ballot(state == kLaneWalk);
        if (ready == 0ull && walk == 0ull) {
            if (__ballot(l.active) == 0ull) break;   // no work left anywhere
            continue;
        }

        // the wave's scene, read again each step (nothing held across steps)
        const SceneView S = const_copy(&bs->S);
        const FrameArgs F = const_copy(&bs->F);
        const FeatureArgs X = FEAT ? const_copy(&bs->X) : FeatureArgs{};
        // ---- the finished queries of the wave, together ----
        if (ready != 0ull && ((uint32_t)__popcll(ready) >= kPersistShadeLanes || walk == 0ull) &&
            state == kLaneReady) {
            if (l.mode == 0u) {
                if (l.b == 0u) n_primary++;
                else n_ext++;
                ScratchSink sink{{}, A.rec, A.threads, tid, 0u};
                VertexResult R;
                shade_vertex<(EXT & kExtSky) != 0u, (EXT & kExtSdf) != 0u, (EXT & kExtPbr) != 0u,
                             (EXT & kExtWater) != 0u>(S, F, l.b, l.o, l.d, h, l.P, R, sink, &X.sky_args);
                const bool cont = R.alive && (l.b + 1u) < F.max_bounces;
                l.pending = sink.pending | (cont ? 0u : kEnded);
                c.set3(Cold::kNo, R.no);
                c.set3(Cold::kNd, R.nd);
                advance<FEAT, (EXT & kExtSky) != 0u, (EXT & kExtWater) != 0u>(F, X, A.rec, A.threads, tid, accum, reset, l,
                                                                                       c);
            } else {
                n_shadow++;
                float t_max = c.get(Cold::kTmax);
                f3 tr = c.get3(Cold::kTr);
                const uint32_t step_w = c.getu(Cold::kStep);   // step | medium << 31
                const uint32_t step = step_w & 0x7FFFFFFFu;
                bool wet = (step_w >> 31) != 0u;
                if (transmit_step<(EXT & kExtSdf) != 0u, (EXT & kExtWater) != 0u>(S, F, h, step, l.o, l.d, t_max, tr,
                                                                                   l.P.time, wet)) {
                    const uint32_t k = l.mode - 1u;
                    if (any_pos(tr)) {
                        const float4* row = A.rec + ((size_t)k * kRecRows) * A.threads + tid;
                        const float wt = row[(size_t)A.threads].w;
                        const float4 a4 = row[2 * (size_t)A.threads], b4 = row[3 * (size_t)A.threads];
                        l.P.L = l.P.L + shadow_value(F, (int)k, xyz(a4), a4.w, xyz(b4), b4.w, k == 1u ? wt : 1.0f,
                                                     (l.pending & (16u << k)) != 0u, tr);
                    }
                    l.pending &= ~(1u << k);
                    advance<FEAT, (EXT & kExtSky) != 0u, (EXT & kExtWater) != 0u>(F, X, A.rec, A.threads, tid, accum, reset, l,
                                                                                       c);
                } else {
                    c.set(Cold::kTmax, t_max);
                    c.set3(Cold::kTr, tr);
                    c.setu(Cold::kStep, (step + 1u) | (wet ? 0x80000000u : 0u));
                }
            }
            if (l.active) {
                state = kLaneWalk;
                if (walk_begin<(EXT & kExtSdf) != 0u>(S, l.o, l.d, w)) {
                    h = final_hit(w.ab, w.tb);
                    state = kLaneReady;
                }
            } else {
                state = kLaneIdle;
            }
        }
        // ---- traversal steps of every lane with a query in progress ----
#pragma unroll
        for (uint32_t k = 0; k < kTailSteps; ++k) {
            if (state == kLaneWalk && walk_step<false, FEAT, (EXT & kExtSdf) != 0u>(S, l.o, l.d, w, stk, tc, l.P.time)) {
                h = final_hit(w.ab, w.tb);
                state = kLaneReady;
            }
        }
    }
    add_stat(&A.stats->primary, n_primary);
    add_stat(&A.stats->extension, n_ext);
    add_stat(&A.stats->shadow, n_shadow);
}

__global__ void k_batch_begin(uint32_t* work, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) work[i] = 0u;
}

__global__ void k_batch_end(uint32_t* frame) {
    if (threadIdx.x == 0) frame[0] = frame[0] + 1u;
}

__global__ void k_batch_set_frame(uint32_t* frame, uint32_t v) {
    if (threadIdx.x == 0) frame[0] = v;
}

}  // namespace

hipError_t batch_set_frame(uint32_t* frame, uint32_t value, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch_set_frame, dim3(1), dim3(64), 0, stream, frame, value);
    return hipGetLastError();
}

// Variant v: 0 plain; 1 + ext, features with the extensions of ext (sky
// lights, SDF clusters, PBR maps, water: any combination).
template <uint32_t DEPTH, uint32_t E = 0u>
hipError_t launch_persistent_ext(uint32_t ext, uint32_t blocks, const BatchArgs& A, hipStream_t stream) {
    if constexpr (E < kExtVariants) {
        if (ext != E) return launch_persistent_ext<DEPTH, E + 1u>(ext, blocks, A, stream);
        hipLaunchKernelGGL((k_persistent<DEPTH, true, E>), dim3(blocks), dim3(kBlock), 0, stream, A);
        return hipSuccess;
    } else {
        return hipErrorInvalidValue;
    }
}

template <uint32_t DEPTH>
hipError_t launch_persistent(uint32_t variant, uint32_t blocks, const BatchArgs& A, hipStream_t stream) {
    if (variant == 0u) {
        hipLaunchKernelGGL((k_persistent<DEPTH, false>), dim3(blocks), dim3(kBlock), 0, stream, A);
        return hipSuccess;
    }
    return launch_persistent_ext<DEPTH>(variant - 1u, blocks, A, stream);
}

hipError_t launch_batch_frame(const BatchArgs& A, const BatchLaunch& L, hipStream_t stream) {
    if (L.blocks == 0u || (size_t)L.blocks * kBlock != A.threads) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(256), 0, stream, A.work, A.nscenes);
    // variant: 0 plain, 1 + ext features with extensions (an extension implies features)
    const uint32_t v = L.features || L.ext ? 1u + (L.ext & (kExtVariants - 1u)) : 0u;
    hipError_t e;
    switch (L.stack_lds) {
        case 16u: e = launch_persistent<16u>(v, L.blocks, A, stream); break;
        case 24u: e = launch_persistent<24u>(v, L.blocks, A, stream); break;
        case 32u: e = launch_persistent<32u>(v, L.blocks, A, stream); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch_end, dim3(1), dim3(64), 0, stream, A.frame);
    return hipGetLastError();
}

}  // namespace dmt
