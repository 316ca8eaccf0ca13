// Streaming wavefront kernels of the gfx950 path tracer (see wavefront.h).
//
//   k_generate    first camera ray of every pixel (8x8 tile order)
//   k_extend      closest hit for the live ray queue (traceScene)
//   k_shade       one vertex per live path: emits shadow queries, the next ray
//                 or a "finished" entry
//   k_shadow<k>   transmittance + contribution of light (k=0) / env (k=1) NEE
//   k_finish      sample bookkeeping of ended paths; the frame value per pixel
//   k_regen       camera rays of the next samples, 8x8 tile order
//   k_accumulate  folds the frame's values into the accumulation buffer
//
// Queue appends are wave-aggregated: one atomic per wave64 for each queue,
// slots from a ballot prefix.  Queue order is not deterministic, the image
// is: every path's arithmetic is done by one lane in a fixed order.
#include "path.h"
#include "tail.h"
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

// Workgroup-aggregated append: every thread of the block calls it (block-
// uniform loops); one global atomic per workgroup and counter.
__device__ __forceinline__ uint32_t block_append(uint32_t* ctr, bool pred, uint32_t* s_cnt) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < kBlock / 64u; ++w) {
            const uint32_t c = s_cnt[w];
            s_cnt[w] = total;
            total += c;
        }
        s_cnt[kBlock / 64u] = total ? atomicAdd(ctr, total) : 0u;
    }
    __syncthreads();
    const uint32_t slot = s_cnt[kBlock / 64u] + s_cnt[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    __syncthreads();   // s_cnt is reused by the next call
    return slot;
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
    return stack_at(lds + threadIdx.x, kBlock, kLdsDepth, spill + (size_t)gid * spill_depth, spill_depth);
}

// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_generate(FrameArgs F, PathState ps, RayQueue q, uint32_t* ctr) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    if (gid == 0) {
        for (uint32_t k = 0; k < kCounterSlots; ++k) ctr[k] = 0u;
        ctr[kCtrRays + 0] = F.width * F.height;
    }
    const uint32_t nslots = camera_tiles(F) * 64u;
    for (uint32_t qs = gid; qs < nslots; qs += stride) camera_path(F, F.frame_index, ps, q, qs);
}

// The tile order of a frame from the per-pixel costs of the frame set's
// previous frame: one wave per 8x8 tile sums its pixels' costs (each capped at
// 2^25 ticks) and files the tile in the log2 bucket of the sum, counted in
// one of kTileBinCopies copies of the bins (same-address global atomics from
// every tile would serialise).
__global__ void __launch_bounds__(kBlock) k_tile_buckets(FrameArgs F, const uint32_t* cost, uint32_t* bins,
                                                         uint32_t* bucket, uint32_t tiles) {
    const uint32_t tile = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63u;
    if (tile >= tiles) return;   // wave-uniform
    const uint32_t tiles_x = (F.width + 7u) / 8u;
    const uint32_t x = (tile % tiles_x) * 8u + (lane & 7u), y = (tile / tiles_x) * 8u + (lane >> 3);
    const uint32_t c = (x < F.width && y < F.height) ? min(cost[y * F.width + x], 1u << 25) : 0u;
    const uint32_t key = wave_sum(c);
    if (lane == 0u) {
        const uint32_t b = key ? 32u - (uint32_t)__clz(key) : 0u;
        bucket[tile] = b;
        atomicAdd(&bins[(blockIdx.x % kTileBinCopies) * kTileBuckets + b], 1u);
    }
}

// Each tile's position: after the tiles of heavier buckets, then within its
// bucket a range its workgroup reserves for the bucket's tiles of its chunk,
// then the tile's rank there (in no particular order).  Every tile is counted
// once, so the positions are 0 .. tiles - 1, each taken once.
__global__ void __launch_bounds__(kBlock) k_tile_order(const uint32_t* bins, uint32_t* cursor, const uint32_t* bucket,
                                                       uint32_t* order, uint32_t tiles) {
    __shared__ uint32_t s_start[kTileBuckets], s_base[kTileBuckets], s_cnt[kTileBuckets];
    if (threadIdx.x < kTileBuckets) {
        uint32_t n = 0u;
        for (uint32_t k = 0; k < kTileBinCopies; ++k) n += bins[k * kTileBuckets + threadIdx.x];
        s_cnt[threadIdx.x] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0u;
        for (int b = (int)kTileBuckets - 1; b >= 0; --b) {
            s_start[b] = acc;
            acc += s_cnt[b];
        }
    }
    __syncthreads();
    if (threadIdx.x < kTileBuckets) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t chunk = (tiles + gridDim.x - 1u) / gridDim.x;
    const uint32_t lo = blockIdx.x * chunk, hi = min(lo + chunk, tiles);
    for (uint32_t t = lo + threadIdx.x; t < hi; t += kBlock) atomicAdd(&s_cnt[bucket[t]], 1u);
    __syncthreads();
    if (threadIdx.x < kTileBuckets) {
        const uint32_t n = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = n ? atomicAdd(&cursor[threadIdx.x], n) : 0u;
        s_cnt[threadIdx.x] = 0u;
    }
    __syncthreads();
    for (uint32_t t = lo + threadIdx.x; t < hi; t += kBlock) {
        const uint32_t b = bucket[t];
        const uint32_t pos = s_start[b] + s_base[b] + atomicAdd(&s_cnt[b], 1u);
        if (pos < tiles) order[pos] = t;
    }
}

template <bool COUNT>
__global__ void __launch_bounds__(kBlock) k_extend(SceneView S, RayQueue q, uint32_t* ctr, uint32_t in_par,
                                                   uint2* hits, uint32_t* spill, uint32_t spill_depth,
                                                   DeviceStats* st) {
    __shared__ uint32_t lds[kLdsDepth * kBlock];
    const uint32_t n = ctr[kCtrRays + in_par];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    if (gid == 0) {
        ctr[kCtrRays + (in_par ^ 1u)] = 0u;     // the queue this iteration fills
        ctr[kCtrExtend + (in_par ^ 1u)] = 0u;   // the next iteration's ray counter
        if (COUNT) atomicAdd(&st->ext_rays, (unsigned long long)n);
    }
    Stack stk = make_stack(lds, spill, spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    // lane i starts on ray i (slots are in 8x8 tile order: a wave starts on
    // one tile), then takes rays stride, stride + 1, ... as its rays finish
    trace_queue<COUNT>(
        S, gid, stride, n, &ctr[kCtrExtend + in_par], stk, tc,
        [&](uint32_t i, f3& o, f3& d) {
            const float4 a = q.o[i], b = q.d[i];
            o = xyz(a);
            d = xyz(b);
        },
        [&](uint32_t i, HitRec h) { hits[i] = make_uint2(__float_as_uint(h.t), h.code); });
    if (COUNT) {
        add_stat(&st->ext_nodes, tc.nodes);
        add_stat(&st->ext_tris, tc.tris);
        add_stat(&st->ext_spheres, tc.spheres);
    }
}

__global__ void __launch_bounds__(kBlock) k_shade(SceneView S, FrameArgs F, uint32_t in_par, RayQueue qin,
                                                  RayQueue qout, uint32_t* ctr, const uint2* hits, PathState ps,
                                                  ShadowQueue sl, ShadowQueue se, uint32_t* ended,
                                                  DeviceStats* st) {
    __shared__ uint32_t s_cnt[kBlock / 64u + 1u];
    const uint32_t n = ctr[kCtrRays + in_par];
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t n_primary = 0, n_ext = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {   // block-uniform
        const uint32_t i = base + threadIdx.x;
        bool cont = false;
        uint32_t pix = 0u;
        VertexResult R;
        if (i < n) {
            const float4 o4 = qin.o[i], d4 = qin.d[i];
            pix = __float_as_uint(d4.w);
            const uint2 hr = hits[i];
            const float4 Tv = ps.T[pix], Lv = ps.L[pix];
            const uint32_t meta = ps.meta[pix];
            const uint32_t b = meta & 0xFFu;
            const uint32_t flags = __float_as_uint(Lv.w);
            PathVertexState P{xyz(Tv), xyz(Lv), Tv.w, (flags & 1u) != 0u, (flags & 2u) != 0u, ps.rng[pix]};
            SlotSink sink{{}, sl, se, ShadowQueue{}, ShadowQueue{}, i, pix, 0u};
            shade_vertex(S, F, b, xyz(o4), xyz(d4), HitRec{__uint_as_float(hr.x), hr.y}, P, R, sink);
            if (!(sink.emitted & 1u)) sl.d[i] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
            if (!(sink.emitted & 2u)) se.d[i] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
            if (b == 0u) n_primary++;
            else n_ext++;
            const uint32_t nflags = (P.skip_sky ? 1u : 0u) | (P.prev_lambert ? 2u : 0u);
            ps.T[pix] = make_float4(P.T_.x, P.T_.y, P.T_.z, P.prev_pdf);
            ps.L[pix] = make_float4(P.L.x, P.L.y, P.L.z, __uint_as_float(nflags));
            ps.rng[pix] = P.seed;
            cont = R.alive && (b + 1u) < F.max_bounces;
            if (cont) ps.meta[pix] = meta + 1u;
            ended[i] = cont ? kMiss : pix;
        }
        const uint32_t slot = block_append(&ctr[kCtrRays + (in_par ^ 1u)], cont, s_cnt);
        if (cont) {
            qout.o[slot] = make_float4(R.no.x, R.no.y, R.no.z, 0.0f);
            qout.d[slot] = make_float4(R.nd.x, R.nd.y, R.nd.z, __uint_as_float(pix));
        }
    }
    add_stat(&st->primary, n_primary);
    add_stat(&st->extension, n_ext);
}

template <int KIND, bool COUNT>
__global__ void __launch_bounds__(kBlock) k_shadow(SceneView S, FrameArgs F, uint32_t in_par, ShadowQueue q,
                                                   const uint32_t* ctr, PathState ps, uint32_t* spill,
                                                   uint32_t spill_depth, DeviceStats* st) {
    __shared__ uint32_t lds[kLdsDepth * kBlock];
    const uint32_t n = ctr[kCtrRays + in_par];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    Stack stk = make_stack(lds, spill, spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    uint32_t steps = 0;
    for (uint32_t i = gid; i < n; i += stride) {
        const float4 d4 = q.d[i];
        const uint32_t tag = __float_as_uint(d4.w);
        if (!(tag & kValidBit)) continue;
        const float4 o4 = q.o[i], a4 = q.a[i], b4 = q.b[i];
        const uint32_t pix = tag & ~(kClampBit | kValidBit);
        const float4 Lv = ps.L[pix];
        f3 L = xyz(Lv);
        if (KIND == 1 && b4.w < 0.0f) {
            L = L + xyz(a4);   // ambient term of the no-env Lambert branch
        } else {
            f3 tr = transmittance<COUNT>(S, F, xyz(o4), xyz(d4), o4.w, stk, tc, steps);
            if (!any_pos(tr)) continue;
            L = L + shadow_value(F, KIND, xyz(a4), a4.w, xyz(b4), b4.w, KIND == 1 ? q.w[i] : 1.0f,
                                 (tag & kClampBit) != 0u, tr);
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

// ---------------------------------------------------------------------------
// k_tail: the rest of the frame (or all of it) in one launch.
//
// A persistent grid takes the paths of the live queue (the camera rays of
// k_generate, or the paths still alive after the per-bounce kernels) and
// runs each to its end, then the pixel's remaining samples, writing the
// pixel's frame value to PathState::F.  Every lane advances its closest-hit
// query one BVH node or leaf per loop step (Walk); a lane whose query is done
// waits until enough of the wave is done too, and the wave then processes
// those lanes together: the vertex (shade_vertex), or one step of a shadow
// query (transmit_step).  The traversal steps keep most lanes busy instead of
// every query waiting for the wave's longest ray, and the shading code runs
// for many lanes at once; idle lanes take new paths from the queue.  Each
// path's arithmetic is the one the per-bounce kernels do, in the same order
// (vertex shading, light-NEE then env-NEE contribution, next vertex; sample
// sums as k_finish forms them), so the pixel values are the same.  The
// scene's data stay kernel arguments, read where they are used.
//
// Five waves per SIMD (the LDS of a workgroup, stack and cold fields, allows
// five per CU): 96 VGPRs, a few spilled values outside the traversal loop.
//
// FEAT (frames with camera-sample features: thin lens, shutter time, hero
// wavelength, a fixed jitter, a translated mesh): there is no queue; entry i
// is slot i of the frame's 8x8 tile order, whose pixel's first sample the
// lane starts itself (camera_sample), and the shutter time and wavelength
// ride in the lane's path state.  EXT (implies FEAT), kExtSky: the frame has
// sky lights (their records and NEE exist only in such a variant); kExtSdf:
// the scene has SDF clusters (likewise their traversal).
template <bool COUNT, bool FEAT, uint32_t EXT = 0u>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(tail_waves(EXT), tail_waves(EXT))))
k_tail(SceneView S, FrameArgs F, FeatureArgs X, uint32_t in_par, RayQueue q, uint32_t* ctr, PathState ps,
       ShadowQueue sl, ShadowQueue se, ShadowQueue s2, ShadowQueue s3, uint32_t* spill, uint32_t spill_depth,
       DeviceStats* st) {
    __shared__ uint32_t lds[kLdsDepth * kBlock];
    __shared__ float cold[tail::TailCold::kRows * kBlock];
    tail::TailCold c{cold + threadIdx.x};
    const uint32_t n = FEAT ? camera_tiles(F) * 64u : ctr[kCtrRays + in_par];
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    Stack stk = make_stack(lds, spill, spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    uint32_t n_primary = 0, n_ext = 0, n_shadow = 0;
    bool drained = false;
    tail::TailLane l{};
    l.active = false;
    // Each lane advances its query one node or leaf per loop step (Walk);
    // a lane whose query is done waits (tail::kLaneReady) until enough of the wave
    // is done too, and the wave then shades those lanes together.  So the
    // traversal steps keep most lanes busy instead of waiting, query by
    // query, for the wave's longest ray, and the shading code runs for many
    // lanes at once.
    Walk w{};
    HitRec h{0.0f, kMiss};
    uint32_t state = tail::kLaneIdle;
    for (;;) {
        // hand queue entries to idle lanes (wave-uniform)
        const unsigned long long idle = __ballot(state == tail::kLaneIdle);
        const unsigned long long walking = __ballot(state == tail::kLaneWalk);
        if (idle != 0ull && !drained && ((uint32_t)__popcll(idle) >= tail::kTailRefillLanes || walking == 0ull)) {
            const uint32_t want = (uint32_t)__popcll(idle);
            uint32_t base = 0u;
            if (idle & (1ull << lane) ? (idle & below) == 0ull : false) base = atomicAdd(&ctr[kCtrTail], want);
            base = __shfl(base, __ffsll((long long)idle) - 1);
            const uint32_t i = base + (uint32_t)__popcll(idle & below);
            if (FEAT && state == tail::kLaneIdle && i < n) {
                const uint32_t tiles_x = (F.width + 7u) / 8u;
                const uint32_t tile = ps.tile_order ? ps.tile_order[i >> 6] : i >> 6, ln = i & 63u;
                const uint32_t x = (tile % tiles_x) * 8u + (ln & 7u), y = (tile / tiles_x) * 8u + (ln >> 3);
                if (x < F.width && y < F.height) {
                    l.pix = y * F.width + x;
                    if (ps.cost) ps.cost[l.pix] = (uint32_t)__builtin_amdgcn_s_memrealtime();
                    l.slot = l.pix;   // the pixel's shadow-record rows
                    l.P = PathVertexState{f3{1.0f, 1.0f, 1.0f}, f3{0.0f, 0.0f, 0.0f}, 0.0f, false, false,
                                          pixel_seed(F, x, y, F.frame_index)};
                    if (EXT & kExtWater) l.P.in_water = X.cam_in_water != 0u;
                    camera_sample(F, X, x, y, 0u, l.P.seed, l.o, l.d, l.P.time, l.P.lambda);
                    l.b = 0u;
                    c.setu(tail::TailCold::kS, 0u);
                    l.mode = 0u;
                    l.active = true;
                    state = tail::kLaneWalk;
                    if (walk_begin<(EXT & kExtSdf) != 0u>(S, l.o, l.d, w)) {
                        h = final_hit(w.ab, w.tb);
                        state = tail::kLaneReady;
                    }
                }
            } else if (!FEAT && state == tail::kLaneIdle && i < n) {
                const float4 o4 = q.o[i], d4 = q.d[i];
                l.pix = __float_as_uint(d4.w);
                if (ps.cost) ps.cost[l.pix] = (uint32_t)__builtin_amdgcn_s_memrealtime();
                l.slot = i;
                l.o = xyz(o4);
                l.d = xyz(d4);
                const float4 Tv = ps.T[l.pix], Lv = ps.L[l.pix];
                const uint32_t meta = ps.meta[l.pix];
                const uint32_t flags = __float_as_uint(Lv.w);
                l.P = PathVertexState{xyz(Tv), xyz(Lv), Tv.w, (flags & 1u) != 0u, (flags & 2u) != 0u, ps.rng[l.pix]};
                l.b = meta & 0xFFu;
                const uint32_t s = (meta >> 8) & 0xFFu;
                c.setu(tail::TailCold::kS, s);
                if (s != 0u) c.set3(tail::TailCold::kSum, xyz(ps.F[l.pix]));
                l.mode = 0u;
                l.active = true;
                state = tail::kLaneWalk;
                if (walk_begin<(EXT & kExtSdf) != 0u>(S, l.o, l.d, w)) {
                    h = final_hit(w.ab, w.tb);
                    state = tail::kLaneReady;
                }
            }
            if (base + want >= n) drained = true;
        }
        const unsigned long long ready = __ballot(state == tail::kLaneReady);
        const unsigned long long walk = __ballot(state == tail::kLaneWalk);
        if (ready == 0ull && walk == 0ull) {
            if (drained) break;
            continue;
        }
        // the finished queries of the wave, together: the vertex, or one step
        // of a shadow query; then the lane's next query
        const uint32_t shade_at = drained ? tail::kTailShadeLanesDrained : tail::kTailShadeLanes;
        if (ready != 0ull && ((uint32_t)__popcll(ready) >= shade_at || walk == 0ull) && state == tail::kLaneReady) {
            if (l.mode == 0u) {
                if (l.b == 0u) n_primary++;
                else n_ext++;
                SlotSink sink{{}, sl, se, s2, s3, l.slot, l.pix, 0u};
                VertexResult R;
                shade_vertex<(EXT & kExtSky) != 0u, (EXT & kExtSdf) != 0u, (EXT & kExtPbr) != 0u,
                             (EXT & kExtWater) != 0u>(S, F, l.b, l.o, l.d, h, l.P, R, sink, &X.sky_args);
                const bool cont = R.alive && (l.b + 1u) < F.max_bounces;
                l.pending = sink.emitted | (cont ? 0u : tail::kEnded);
                if (cont) {
                    c.set3(tail::TailCold::kNo, R.no);
                    c.set3(tail::TailCold::kNd, R.nd);
                }
                tail::tail_advance<FEAT, (EXT & kExtSky) != 0u, (EXT & kExtWater) != 0u>(F, X, sl, se, s2, s3, ps, l,
                                                                                                    c);
            } else {
                n_shadow++;
                float t_max = c.get(tail::TailCold::kTmax);
                f3 tr = c.get3(tail::TailCold::kTr);
                const uint32_t step_w = c.getu(tail::TailCold::kStep);   // step | medium << 31
                const uint32_t step = step_w & 0x7FFFFFFFu;
                bool wet = (step_w >> 31) != 0u;
                if (transmit_step<(EXT & kExtSdf) != 0u, (EXT & kExtWater) != 0u>(S, F, h, step, l.o, l.d, t_max, tr,
                                                                                   l.P.time, wet)) {
                    const uint32_t k = l.mode - 1u;
                    if (any_pos(tr)) {
                        // fields picked one by one: a selected reference to one of
                        // the queue arguments would put them in private memory
                        const float4* qa = k == 0u ? sl.a : se.a;
                        const float4* qb = k == 0u ? sl.b : se.b;
                        const float4* qd = k == 0u ? sl.d : se.d;
                        if ((EXT & kExtSky) && k >= 2u) {
                            qa = k == 2u ? s2.a : s3.a;
                            qb = k == 2u ? s2.b : s3.b;
                            qd = k == 2u ? s2.d : s3.d;
                        }
                        const float4 a4 = qa[l.slot], b4 = qb[l.slot];
                        const uint32_t tag = __float_as_uint(qd[l.slot].w);
                        l.P.L = l.P.L + shadow_value(F, (int)k, xyz(a4), a4.w, xyz(b4), b4.w,
                                                     k == 1u ? se.w[l.slot] : 1.0f, (tag & kClampBit) != 0u, tr);
                    }
                    l.pending &= ~(1u << k);
                    tail::tail_advance<FEAT, (EXT & kExtSky) != 0u, (EXT & kExtWater) != 0u>(F, X, sl, se, s2, s3, ps, l,
                                                                                                    c);
                } else {   // transmit_step ends every query by its step refract_bounces
                    c.set(tail::TailCold::kTmax, t_max);
                    c.set3(tail::TailCold::kTr, tr);
                    c.setu(tail::TailCold::kStep, (step + 1u) | (wet ? 0x80000000u : 0u));
                }
            }
            if (l.active) {
                state = tail::kLaneWalk;
                if (walk_begin<(EXT & kExtSdf) != 0u>(S, l.o, l.d, w)) {
                    h = final_hit(w.ab, w.tb);
                    state = tail::kLaneReady;
                }
            } else {
                state = tail::kLaneIdle;
            }
        }
        // traversal steps of every lane with a query in progress.  With SDF
        // clusters a pass is either a node pass (lanes at BVH nodes and
        // leaves) or a trace pass (lanes inside a cluster's sphere trace):
        // the wave runs one of the two codes per pass, the trace pass once
        // kSdfGather lanes trace or nothing else walks.
        if constexpr ((EXT & kExtSdf) != 0u && kSdfGather != 0u) {
            const unsigned long long tracing = __ballot(state == tail::kLaneWalk && w.sdf_ci != kMiss);
            const unsigned long long nodes = __ballot(state == tail::kLaneWalk && w.sdf_ci == kMiss);
            const bool trace_pass = tracing != 0ull && ((uint32_t)__popcll(tracing) >= kSdfGather || nodes == 0ull);
            const uint32_t pass_steps = trace_pass ? kSdfBurst : kTailSteps;
            for (uint32_t k = 0; k < pass_steps; ++k) {
                if (state == tail::kLaneWalk && (w.sdf_ci != kMiss) == trace_pass &&
                    walk_step<COUNT, FEAT, true>(S, l.o, l.d, w, stk, tc, l.P.time)) {
                    h = final_hit(w.ab, w.tb);
                    state = tail::kLaneReady;
                }
            }
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kTailSteps; ++k) {
                if (state == tail::kLaneWalk &&
                    walk_step<COUNT, FEAT, (EXT & kExtSdf) != 0u>(S, l.o, l.d, w, stk, tc, l.P.time)) {
                    h = final_hit(w.ab, w.tb);
                    state = tail::kLaneReady;
                }
            }
        }
    }
    tail::add_stat(&st->primary, n_primary);
    tail::add_stat(&st->extension, n_ext);
    tail::add_stat(&st->shadow, n_shadow);
    if (COUNT) {
        tail::add_stat(&st->tail_nodes, tc.nodes);
        tail::add_stat(&st->tail_tris, tc.tris);
        tail::add_stat(&st->tail_spheres, tc.spheres);
    }
}

// Sample bookkeeping of the paths that ended this iteration: the next sample
// of the pixel is flagged (k_regen issues it in 8x8 tile order so the new
// waves stay coherent); after the last sample the frame's value goes to ps.F
// for k_accumulate.
__global__ void __launch_bounds__(kBlock) k_finish(FrameArgs F, uint32_t in_par, PathState ps, const uint32_t* ctr,
                                                   const uint32_t* ended, uint8_t* regen) {
    const uint32_t n = ctr[kCtrRays + in_par];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = gid; i < n; i += stride) {
        const uint32_t p = ended[i];
        if (p == kMiss) continue;
        const uint32_t s = (ps.meta[p] >> 8) & 0xFFu;
        const f3 L = xyz(ps.L[p]);
        const f3 sum = (s == 0u) ? L : xyz(ps.F[p]) + L;
        if (s + 1u < F.spp) {
            ps.F[p] = make_float4(sum.x, sum.y, sum.z, 0.0f);
            ps.meta[p] = meta_pack(0u, s + 1u, 0u);   // the next sample continues the pixel's stream
            regen[p] = 1u;
        } else {
            const f3 v = frame_value(F, sum);
            ps.F[p] = make_float4(v.x, v.y, v.z, 0.0f);
        }
    }
}

// Camera rays of the flagged pixels, appended in 8x8 tile order.
__global__ void __launch_bounds__(kBlock) k_regen(FrameArgs F, uint32_t out_par, PathState ps, uint32_t* ctr,
                                                  uint8_t* regen, RayQueue qout) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t tiles_x = (F.width + 7u) / 8u;
    const uint32_t tiles_y = (F.height + 7u) / 8u;
    const uint32_t nslots = tiles_x * tiles_y * 64u;
    __shared__ uint32_t s_cnt[kBlock / 64u + 1u];
    // block-uniform trip count
    for (uint32_t q0 = blockIdx.x * blockDim.x; q0 < nslots; q0 += stride) {
        const uint32_t qs = q0 + threadIdx.x;
        const uint32_t tile = qs >> 6, lane = qs & 63u;
        const uint32_t x = (tile % tiles_x) * 8u + (lane & 7u), y = (tile / tiles_x) * 8u + (lane >> 3);
        const bool inside = qs < nslots && x < F.width && y < F.height;
        const uint32_t p = y * F.width + x;
        const bool go = inside && regen[p] != 0u;
        const uint32_t slot = block_append(&ctr[kCtrRays + out_par], go, s_cnt);
        if (go) {
            regen[p] = 0u;
            uint32_t seed = ps.rng[p];
            f3 rd = camera_ray(F, x, y, seed);
            start_path(ps, qout, slot, p, F.pos, rd, seed, ps.meta[p]);
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_accumulate(FrameArgs F, PathState ps, float4* accum) {
    const uint32_t npix = F.width * F.height;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = gid; p < npix; p += stride) {
        const f3 v = xyz(ps.F[p]);
        const bool reset = F.reset_accum != 0u;
        accum[p] = accumulate(F, reset ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : accum[p], reset, v);
    }
}

__global__ void k_snapshot(const uint32_t* ctr, uint32_t par, uint32_t* host_slot) {
    if (threadIdx.x == 0) *host_slot = ctr[kCtrRays + par];
}

__global__ void k_reset_counters(uint32_t* ctr) {
    for (uint32_t k = threadIdx.x; k < kCounterSlotsAll; k += blockDim.x) ctr[k] = 0u;
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
        HitRec h = trace_closest<false, false, true>(S, o, d, stk, tc);
        HitOut r{};
        if (h.code != kMiss) {
            Surface sf = surface_at<true>(S, o, d, h);
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

// clip = M * (p, 1) with M column-major (mul(M, v) of PathTrace.slang:5517-5518
// on the engine's glm::value_ptr layout, Engine.cpp:8258-8259).
__device__ __forceinline__ float4 project(const float* m, f3 p) {
    return make_float4(((m[0] * p.x + m[4] * p.y) + m[8] * p.z) + m[12], ((m[1] * p.x + m[5] * p.y) + m[9] * p.z) + m[13],
                       ((m[2] * p.x + m[6] * p.y) + m[10] * p.z) + m[14],
                       ((m[3] * p.x + m[7] * p.y) + m[11] * p.z) + m[15]);
}

__device__ __forceinline__ float2 to_pixel(float4 c, float w, float h) {
    const float cw = fmaxf(c.w, 1e-6f);
    return make_float2(((c.x / cw) * 0.5f + 0.5f) * w, ((-c.y / cw) * 0.5f + 0.5f) * h);
}

// One pixel-centre ray per thread (grid-stride): the G-buffer writes of
// PathTrace.slang:5463-5648 with the sky sentinels of a miss.
__global__ void __launch_bounds__(kBlock) k_gbuffer(SceneView S, GbufferArgs A, uint32_t* spill, uint32_t spill_depth) {
    __shared__ uint32_t lds[kLdsDepth * kBlock];
    const FrameArgs& F = A.F;
    const uint32_t n = F.width * F.height;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    Stack stk = make_stack(lds, spill, spill_depth);
    TraceCounters tc{0u, 0u, 0u};
    const f3 fwdn = normalize(F.fwd);
    const float W = (float)F.width, H = (float)F.height;
    for (uint32_t p = gid; p < n; p += stride) {
        const uint32_t x = p % F.width, y = p / F.width;
        const float ux = (((float)x + 0.5f) / W) * 2.0f - 1.0f;
        float uy = (((float)y + 0.5f) / H) * 2.0f - 1.0f;
        uy = -uy;
        const f3 d = normalize((F.fwd + F.right * ((ux * F.aspect) * F.fovtan)) + F.up * (uy * F.fovtan));
        // shutter time 1: a translated mesh at its current translation
        // (PathTrace.slang:5470-5478)
        const HitRec h = trace_closest<false, true, true>(S, F.pos, d, stk, tc, 1.0f);
        f3 world, nrm{0.0f, 0.0f, 0.0f}, alb{0.0f, 0.0f, 0.0f}, f0{0.0f, 0.0f, 0.0f};
        float depth, rough = 1.0f, dist = 0.0f;
        if (h.code != kMiss) {
            const Surface sf = surface_at<true>(S, F.pos, d, h);
            world = F.pos + d * h.t;
            nrm = sf.n;
            alb = sf.albedo;
            depth = h.t * dot(d, fwdn);
            const float r = fminf(fmaxf(sf.rough, 0.0f), 1.0f);
            if (sf.mat == 1u) f0 = sf.albedo;                            // metal
            else if (sf.mat == 2u) f0 = f3{0.04f, 0.04f, 0.04f};         // dielectric
            rough = sf.mat == 0u ? 1.0f : r;                             // Lambert: no specular lobe
            dist = h.t * (1.0f - r);
        } else {
            world = F.pos + d * 1000.0f;   // far point: the sky still moves
            depth = 1.0e10f;
        }
        const float2 cur = to_pixel(project(A.vp_curr, world), W, H);
        const float2 pre = to_pixel(project(A.vp_prev, world), W, H);
        float* const* o = A.out;
        const size_t b = (size_t)p * A.stride;
        if (o[kGbDepth]) o[kGbDepth][b] = depth;
        if (o[kGbMotion]) {
            o[kGbMotion][b] = pre.x - cur.x;
            o[kGbMotion][b + 1] = pre.y - cur.y;
        }
        const f3 v4[3] = {nrm, alb, f0};
        for (int k = 0; k < 3; ++k) {
            float* q = o[kGbNormal + k];
            if (!q) continue;
            q[b] = v4[k].x;
            q[b + 1] = v4[k].y;
            q[b + 2] = v4[k].z;
            q[b + 3] = 0.0f;
        }
        if (o[kGbRoughness]) o[kGbRoughness][b] = rough;
        if (o[kGbSpecularDistance]) o[kGbSpecularDistance][b] = dist;
        if (o[kGbCloudTrans]) o[kGbCloudTrans][b] = 1.0f;   // no clouds: frame_cloud_trans stays 1
    }
}

// ---------------------------------------------------------------------------
// host side

FrameArgs make_frame_args(const float pf[4], const float fa[4], const float rt[4], const float up[4], uint32_t width,
                          uint32_t height, uint32_t frame_index, uint32_t spp, uint32_t max_bounces,
                          uint32_t reset_accum, uint32_t mis, uint32_t ltree, uint32_t refract, float clamp,
                          float ema, uint32_t flags) {
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
    F.rr_enabled = (flags & 1u) ? 0u : 1u;   // DMT_FRAME_NO_RUSSIAN_ROULETTE
    F.cauchy_c = 0.0f;
    return F;
}

uint32_t grid_for(uint32_t n) {
    uint32_t g = (n + kBlock - 1u) / kBlock;
    if (g < 1u) g = 1u;
    return g < kMaxGridBlocks ? g : kMaxGridBlocks;
}

template <bool COUNT>
static hipError_t iteration(const SceneView& S, const FrameArgs& F, Wavefront& W, hipStream_t stream, uint32_t par,
                            uint32_t g, LaunchOptions& o) {
    const bool ev = o.extend_events && o.events_used + 2u <= o.max_events;
    if (ev) {
        o.event_kinds[o.events_used / 2u] = 0u;
        (void)hipEventRecord(o.extend_events[o.events_used], stream);
    }
    const uint32_t ge = o.extend_blocks ? (o.extend_blocks < g ? o.extend_blocks : g) : g;
    hipLaunchKernelGGL(k_extend<COUNT>, dim3(ge), dim3(kBlock), 0, stream, S, W.rq[par], W.counters, par, W.hits,
                       W.spill, W.spill_depth, W.stats);
    if (ev) {
        (void)hipEventRecord(o.extend_events[o.events_used + 1u], stream);
        o.events_used += 2u;
    }
    hipLaunchKernelGGL(k_shade, dim3(g), dim3(kBlock), 0, stream, S, F, par, W.rq[par], W.rq[par ^ 1u], W.counters,
                       W.hits, W.ps, W.sq[0], W.sq[1], W.finished, W.stats);
    hipLaunchKernelGGL((k_shadow<0, COUNT>), dim3(g), dim3(kBlock), 0, stream, S, F, par, W.sq[0], W.counters,
                       W.ps, W.spill, W.spill_depth, W.stats);
    hipLaunchKernelGGL((k_shadow<1, COUNT>), dim3(g), dim3(kBlock), 0, stream, S, F, par, W.sq[1], W.counters,
                       W.ps, W.spill, W.spill_depth, W.stats);
    hipLaunchKernelGGL(k_finish, dim3(g), dim3(kBlock), 0, stream, F, par, W.ps, W.counters, W.finished, W.regen);
    if (F.spp > 1u)
        hipLaunchKernelGGL(k_regen, dim3(g), dim3(kBlock), 0, stream, F, par ^ 1u, W.ps, W.counters, W.regen,
                           W.rq[par ^ 1u]);
    return hipGetLastError();
}

// The feature variant of k_tail for the frame's extension bits (one
// instantiation per combination).
template <bool COUNT, uint32_t E = 0u, class Go>
void launch_tail_variant(uint32_t ext, Go&& go) {
    if constexpr (E < kExtVariants) {
        if (ext == E) go(k_tail<COUNT, true, E>);
        else launch_tail_variant<COUNT, E + 1u>(ext, go);
    }
}

hipError_t launch_frame(const SceneView& S, const FrameArgs& F, Wavefront& W, float4* accum, hipStream_t stream,
                        LaunchOptions& o) {
    const uint32_t npix = F.width * F.height;
    const uint32_t g = grid_for(npix);
    o.events_used = 0;
    o.iterations = 0;
    const uint32_t tg = o.tail_blocks ? (o.tail_blocks < g ? o.tail_blocks : g) : g;
    // With several samples per pixel a lane runs a pixel's samples one after
    // the other, and the frame's last pixels decide when it ends: tiles then
    // go heaviest first, by the elapsed times the frame set's previous frame
    // recorded, so that a frame alone on the GPU ends with short pixels and
    // fewer idle SIMDs (the many-light scene's lone frame: 17.2 -> 15.1 ms).
    // Which lane runs a pixel changes; what it computes does not.  At one
    // sample per pixel the scattered order costs more cache locality than it
    // saves (the 1080p frame: -5 %), so those frames keep row order and record
    // nothing.  Queued frames need whole tiles.
    PathState ps = W.ps;
    ps.cost = nullptr;
    const uint32_t tiles = ((F.width + 7u) / 8u) * ((F.height + 7u) / 8u);
    if (F.spp > 1u && W.ps.cost && W.tile_order && npix <= W.capacity && tiles <= tile_capacity(W.capacity) &&
        (o.features || (F.width % 8u == 0u && F.height % 8u == 0u))) {
        hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(kBlock), 0, stream, W.counters);
        const uint32_t go = min((tiles + kBlock - 1u) / kBlock, kTileOrderBlocks);
        hipLaunchKernelGGL(k_tile_buckets, dim3((tiles + 3u) / 4u), dim3(kBlock), 0, stream, F, W.ps.cost,
                           W.counters + kCtrTileBins, W.tile_bucket, tiles);
        hipLaunchKernelGGL(k_tile_order, dim3(go), dim3(kBlock), 0, stream, W.counters + kCtrTileBins,
                           W.counters + kCtrTileCursor, W.tile_bucket, W.tile_order, tiles);
        ps.tile_order = W.tile_order;
        ps.cost = W.ps.cost;
    }
    if (o.features) {
        // the whole frame in one k_tail launch that starts the camera paths itself
        if (!ps.tile_order) hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(kBlock), 0, stream, W.counters);
        const bool ev = o.extend_events && o.events_used + 2u <= o.max_events;
        if (ev) {
            o.event_kinds[0] = 1u;
            (void)hipEventRecord(o.extend_events[0], stream);
        }
        const FeatureArgs& X = *o.features;
        const ShadowQueue* sq = W.sq;
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(tg), dim3(kBlock), 0, stream, S, F, X, 0u, W.rq[0], W.counters, ps,
                               sq[0], sq[1], sq[2], sq[3], W.spill, W.spill_depth, W.stats);
        };
        if (o.counting) launch_tail_variant<true>(X.ext & (kExtVariants - 1u), go);
        else launch_tail_variant<false>(X.ext & (kExtVariants - 1u), go);
        if (ev) {
            (void)hipEventRecord(o.extend_events[1], stream);
            o.events_used = 2u;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        o.iterations = 1;
        if (o.accum_after) (void)hipStreamWaitEvent(stream, o.accum_after, 0);
        hipLaunchKernelGGL(k_accumulate, dim3(g), dim3(kBlock), 0, stream, F, W.ps, accum);
        return hipGetLastError();
    }
    const FeatureArgs none{};
    hipLaunchKernelGGL(k_generate, dim3(g), dim3(kBlock), 0, stream, F, ps, W.rq[0], W.counters);
    // Every path ends after at most max_bounces vertices and a pixel's samples
    // run in order, so the frame needs at most spp * max_bounces iterations.
    // Short frames are launched in full; long ones stop once a snapshot of the
    // live queue reads zero.
    const uint64_t bound = (uint64_t)F.spp * (F.max_bounces ? F.max_bounces : 1u);
    const bool poll = bound > 24u && o.poll_events && W.host_count;
    uint32_t snap = 0;
    for (uint64_t it = 0; it < bound; ++it) {
        const uint32_t par = (uint32_t)(it & 1u);
        if (it == o.tail_after) {
            // the live queue (and the pixels' later samples) in one launch
            const bool ev = o.extend_events && o.events_used + 2u <= o.max_events;
            if (ev) {
                o.event_kinds[o.events_used / 2u] = 1u;
                (void)hipEventRecord(o.extend_events[o.events_used], stream);
            }
            if (o.counting)
                hipLaunchKernelGGL((k_tail<true, false>), dim3(tg), dim3(kBlock), 0, stream, S, F, none, par, W.rq[par],
                                   W.counters, ps, W.sq[0], W.sq[1], W.sq[2], W.sq[3], W.spill, W.spill_depth,
                                   W.stats);
            else
                hipLaunchKernelGGL((k_tail<false, false>), dim3(tg), dim3(kBlock), 0, stream, S, F, none, par, W.rq[par],
                                   W.counters, ps, W.sq[0], W.sq[1], W.sq[2], W.sq[3], W.spill, W.spill_depth,
                                   W.stats);
            if (ev) {
                (void)hipEventRecord(o.extend_events[o.events_used + 1u], stream);
                o.events_used += 2u;
            }
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            o.iterations++;
            break;
        }
        hipError_t e = o.counting ? iteration<true>(S, F, W, stream, par, g, o)
                                  : iteration<false>(S, F, W, stream, par, g, o);
        if (e != hipSuccess) return e;
        o.iterations++;
        if (poll && (it + 1u) % kPollChunk == 0u) {
            const uint32_t slot = snap % kPollSlots;
            hipLaunchKernelGGL(k_snapshot, dim3(1), dim3(64), 0, stream, W.counters, par ^ 1u, W.host_count + slot);
            (void)hipEventRecord(o.poll_events[slot], stream);
            ++snap;
            // keep two snapshots in flight; test the older one
            if (snap >= 2u) {
                const uint32_t old = (snap - 2u) % kPollSlots;
                e = hipEventSynchronize(o.poll_events[old]);
                if (e != hipSuccess) return e;
                if (__atomic_load_n(W.host_count + old, __ATOMIC_ACQUIRE) == 0u) break;
            }
        }
    }
    if (o.accum_after) (void)hipStreamWaitEvent(stream, o.accum_after, 0);
    hipLaunchKernelGGL(k_accumulate, dim3(g), dim3(kBlock), 0, stream, F, W.ps, accum);
    return hipGetLastError();
}

hipError_t launch_gbuffer(const SceneView& S, const GbufferArgs& A, Wavefront& W, hipStream_t stream) {
    const uint32_t n = A.F.width * A.F.height;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gbuffer, dim3(grid_for(n)), dim3(kBlock), 0, stream, S, A, W.spill, W.spill_depth);
    return hipGetLastError();
}

hipError_t launch_trace_query(const SceneView& S, const float4* rays, uint32_t n, HitOut* out, Wavefront& W,
                              hipStream_t stream) {
    return launch_trace_query_spill(S, rays, n, out, W.spill, W.spill_depth, kMaxGridBlocks, stream);
}

hipError_t launch_trace_query_spill(const SceneView& S, const float4* rays, uint32_t n, HitOut* out, uint32_t* spill,
                                    uint32_t spill_depth, uint32_t max_blocks, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    uint32_t g = grid_for(n);
    if (g > max_blocks) g = max_blocks;
    hipLaunchKernelGGL(k_trace_query, dim3(g), dim3(kBlock), 0, stream, S, rays, n, out, spill, spill_depth);
    return hipGetLastError();
}

}  // namespace dmt
