// Pieces of the frame-completion kernel (k_tail) and of the camera-ray
// generation shared with the per-bounce kernels: the shadow-record sink, the
// camera path of a pixel, and the lane state of k_tail with its per-vertex
// bookkeeping (tail_advance).
#pragma once
#include "path.h"
#include "wavefront.h"

namespace dmt {

__device__ __forceinline__ void start_path(PathState& ps, RayQueue& q, uint32_t slot, uint32_t p, f3 pos, f3 rd,
                                           uint32_t seed, uint32_t meta) {
    q.o[slot] = make_float4(pos.x, pos.y, pos.z, 0.0f);
    q.d[slot] = make_float4(rd.x, rd.y, rd.z, __uint_as_float(p));
    ps.T[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    ps.L[p] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
    ps.rng[p] = seed;
    ps.meta[p] = meta;
}

// 8x8 tiles of a frame (rows of ceil(width / 8) tiles).
__device__ __forceinline__ uint32_t camera_tiles(const FrameArgs& F) {
    return ((F.width + 7u) / 8u) * ((F.height + 7u) / 8u);
}

// The first path of queue entry qs: lane qs % 64 of 8x8 tile qs / 64 (of
// tile ps.tile_order[qs / 64] when the frame is ordered, which only frames of
// whole tiles are), queued at slot = pixels of the full tile rows above + full
// tiles to the left in this tile row + position inside the (possibly clipped)
// tile, or at qs when ordered: a wave64 traces one tile of neighbouring camera
// rays; the path state stays indexed by pixel.
__device__ __forceinline__ void camera_path(const FrameArgs& F, uint32_t frame_index, PathState& ps, RayQueue& q,
                                            uint32_t qs) {
    const uint32_t tile = ps.tile_order ? ps.tile_order[qs >> 6] : qs >> 6, lane = qs & 63u;
    const uint32_t tiles_x = (F.width + 7u) / 8u;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const uint32_t wx = min(8u, F.width - tx * 8u), hy = min(8u, F.height - ty * 8u);
    const uint32_t lx = lane & 7u, ly = lane >> 3;
    if (lx >= wx || ly >= hy) return;
    const uint32_t x = tx * 8u + lx, y = ty * 8u + ly;
    const uint32_t slot = ps.tile_order ? qs : 8u * ty * F.width + tx * 8u * hy + ly * wx + lx;
    const uint32_t p = y * F.width + x;
    uint32_t seed = pixel_seed(F, x, y, frame_index);
    f3 rd = camera_ray(F, x, y, seed);
    start_path(ps, q, slot, p, F.pos, rd, seed, meta_pack(0u, 0u, 0u));
}

// Writes a path's shadow queries at its own slot of the iteration: light
// NEE, env NEE (or the no-env ambient term), and in frames with sky lights
// the sun or HDRI cluster NEE and the moon NEE.
struct SlotSink : NoGradientHooks {
    ShadowQueue q0, q1, q2, q3;
    uint32_t slot, pix;
    uint32_t emitted;     // bit k: a record of kind k
    __device__ __forceinline__ void emit(int kind, const ShadowRec& r, PathVertexState&) {
        ShadowQueue& s = kind == 0 ? q0 : (kind == 1 ? q1 : (kind == 2 ? q2 : q3));
        const uint32_t tag = pix | kValidBit | (r.clamp ? kClampBit : 0u);
        s.o[slot] = make_float4(r.o.x, r.o.y, r.o.z, r.t_max);
        s.d[slot] = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(tag));
        s.a[slot] = make_float4(r.pre.x, r.pre.y, r.pre.z, r.ndl);
        s.b[slot] = make_float4(r.brdf.x, r.brdf.y, r.brdf.z, r.pdf);
        if (kind == 1) s.w[slot] = r.w;
        emitted |= 1u << kind;
    }
};


namespace tail {


constexpr uint32_t kEnded = 16u;   // pending bit: the path ends after this vertex (bits 0-3: records)

// k_tail lane states
constexpr uint32_t kLaneIdle = 0u;    // no path
constexpr uint32_t kLaneWalk = 1u;    // a closest-hit query in progress
constexpr uint32_t kLaneReady = 2u;   // the query is done; its hit waits for the wave's shading step

// Measured with three traversal steps per pass: shading at 40 ready lanes
// runs the 1.51 M-triangle 1080p frame 3.5 % faster than at 32 (the many-light
// scene within noise); 16 and 24 are slower on both, 48 costs the many-light
// scene 10 %; refill thresholds of 8 and 24 land within noise of 16.
#ifndef DMT_TAIL_SHADE_LANES
#define DMT_TAIL_SHADE_LANES 40u
#endif
#ifndef DMT_TAIL_REFILL_LANES
#define DMT_TAIL_REFILL_LANES 16u
#endif
// Once the queue is empty a wave's lanes only run out, and waiting for 40
// ready lanes stretches the frame's last paths: 12 shortens a lone frame's
// launch by 5 % on both scenes (4: -8 % on the 1080p frame, +0 % many-light;
// 24: -1 % / -4 %).
#ifndef DMT_TAIL_SHADE_LANES_DRAINED
#define DMT_TAIL_SHADE_LANES_DRAINED 12u
#endif
constexpr uint32_t kTailShadeLanes = DMT_TAIL_SHADE_LANES;    // shade once this many lanes are ready
constexpr uint32_t kTailShadeLanesDrained = DMT_TAIL_SHADE_LANES_DRAINED;   // the same once the queue is empty
constexpr uint32_t kTailRefillLanes = DMT_TAIL_REFILL_LANES;  // take new paths once this many are idle

struct TailLane {
    bool active;
    uint32_t mode;        // 0 extension ray, 1 light-NEE shadow ray, 2 env-NEE shadow ray
    f3 o, d;              // current ray
    PathVertexState P;
    uint32_t b;           // bounce
    uint32_t pending;     // shadow records of the vertex (bits 0, 1), kEnded
    uint32_t slot;        // queue entry: index of the lane's shadow records
    uint32_t pix;
};

// Fields of a lane that are not touched at every query: an LDS column per
// thread, field f at base[f * kBlock].
struct TailCold {
    float* base;
    enum : uint32_t { kSum = 0, kS = 3, kTr = 4, kTmax = 7, kStep = 8, kNo = 9, kNd = 12, kRows = 15 };
    __device__ __forceinline__ float get(uint32_t f) const { return base[f * kBlock]; }
    __device__ __forceinline__ void set(uint32_t f, float v) { base[f * kBlock] = v; }
    __device__ __forceinline__ uint32_t getu(uint32_t f) const { return __float_as_uint(get(f)); }
    __device__ __forceinline__ void setu(uint32_t f, uint32_t v) { set(f, __uint_as_float(v)); }
    __device__ __forceinline__ f3 get3(uint32_t f) const { return f3{get(f), get(f + 1), get(f + 2)}; }
    __device__ __forceinline__ void set3(uint32_t f, f3 v) { set(f, v.x); set(f + 1, v.y); set(f + 2, v.z); }
};

template <bool WATER = false>
__device__ __forceinline__ void tail_start_shadow(const ShadowQueue& q, uint32_t k, TailLane& l, TailCold& c) {
    const float4 o4 = q.o[l.slot], d4 = q.d[l.slot];
    l.o = xyz(o4);
    l.d = xyz(d4);
    c.set(TailCold::kTmax, o4.w);
    c.set3(TailCold::kTr, f3{1.0f, 1.0f, 1.0f});
    // the query starts in the path's medium (a Lambert vertex, whose NEE
    // records these are, does not change it): bit 31 of the step field
    c.setu(TailCold::kStep, (WATER && l.P.in_water) ? 0x80000000u : 0u);
    l.mode = k + 1u;
}

// After a vertex or one of its shadow queries: the next shadow query in the
// order the vertex's terms are summed (light, then env NEE and the HDRI
// cluster NEE with an env map; light, sun, moon, then the ambient term
// without one), the next vertex, the next sample of the pixel, or the
// pixel's frame value.  FEAT: the frame's camera-sample features and sky
// lights X are on; the samples draw them (camera_sample) and add their
// radiance weighted by the hero wavelength's response.  SKY: the frame has
// sky lights, whose records (2 and 3) exist only then.  WATER: the frame has
// water (the paths' medium).
template <bool FEAT, bool SKY = false, bool WATER = false>
__device__ __forceinline__ void tail_advance(const FrameArgs& F, const FeatureArgs& X, const ShadowQueue& sl,
                                             const ShadowQueue& se, const ShadowQueue& s2, const ShadowQueue& s3,
                                             const PathState& ps, TailLane& l, TailCold& c) {
    if (l.pending & 1u) {
        tail_start_shadow<WATER>(sl, 0u, l, c);
        return;
    }
    if (!SKY) {
        if (l.pending & 2u) {
            if (se.b[l.slot].w < 0.0f) {
                // ambient term of the no-env Lambert branch: added without a query
                l.P.L = l.P.L + xyz(se.a[l.slot]);
                l.pending &= ~2u;
            } else {
                tail_start_shadow<WATER>(se, 1u, l, c);
                return;
            }
        }
    } else {
        bool ambient = false;
        if (l.pending & 2u) {
            ambient = se.b[l.slot].w < 0.0f;   // the no-env ambient term: after the sun and moon
            if (!ambient) {
                tail_start_shadow<WATER>(se, 1u, l, c);
                return;
            }
        }
        if (l.pending & 4u) {
            tail_start_shadow<WATER>(s2, 2u, l, c);
            return;
        }
        if (l.pending & 8u) {
            tail_start_shadow<WATER>(s3, 3u, l, c);
            return;
        }
        if (ambient) {
            l.P.L = l.P.L + xyz(se.a[l.slot]);
            l.pending &= ~2u;
        }
    }
    if (!(l.pending & kEnded)) {
        l.o = c.get3(TailCold::kNo);
        l.d = c.get3(TailCold::kNd);
        l.b++;
        l.mode = 0u;
        return;
    }
    const uint32_t s = c.getu(TailCold::kS);
    f3 sum;
    if (FEAT) {
        const f3 v = (l.P.lambda >= 0.0f) ? l.P.L * spectral_weight(l.P.lambda) : l.P.L;
        sum = ((s == 0u) ? f3{0.0f, 0.0f, 0.0f} : c.get3(TailCold::kSum)) + v;
    } else {
        sum = (s == 0u) ? l.P.L : c.get3(TailCold::kSum) + l.P.L;
    }
    if (s + 1u < F.spp) {
        // the next sample continues the pixel's stream (k_finish + k_regen)
        c.set3(TailCold::kSum, sum);
        c.setu(TailCold::kS, s + 1u);
        const uint32_t x = l.pix % F.width, y = l.pix / F.width;
        if (FEAT) {
            camera_sample(F, X, x, y, s + 1u, l.P.seed, l.o, l.d, l.P.time, l.P.lambda);
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
        return;
    }
    const f3 v = frame_value(F, sum, FEAT && X.spectral != 0u);
    ps.F[l.pix] = make_float4(v.x, v.y, v.z, 0.0f);
    if (ps.cost) ps.cost[l.pix] = (uint32_t)__builtin_amdgcn_s_memrealtime() - ps.cost[l.pix];
    l.active = false;
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

}  // namespace tail

}  // namespace dmt
